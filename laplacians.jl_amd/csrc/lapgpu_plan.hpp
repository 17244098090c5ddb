// lapgpu_plan.hpp — host-side preprocessing of la and LDLinv into the
// level-ordered device layout used by the sweeps and the SpMV.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lapgpu {

// Laplacian in CSR (== CSC, symmetric), vertex order of the caller.
struct HostCsr {
  int64_t n = 0;
  std::vector<int32_t> ptr;  // n+1 (nnz < 2^31 enforced)
  std::vector<int32_t> idx;
  std::vector<double> val;
};

// Everything the device needs, in "level order": vertices sorted by the
// forward level of the LDLinv column DAG (ties by original index).
struct Plan {
  int64_t n = 0;
  int64_t nnzF = 0;
  std::vector<int32_t> perm;   // perm[pos] = original vertex
  std::vector<int32_t> iperm;  // iperm[vertex] = pos

  // la permuted: row pos -> (pos', value)
  std::vector<int32_t> Lptr, Lidx;
  std::vector<double> Lval;

  // forward sweep, pull form: y0[v] = r[v] + sum_k Fw[k] * y0[Fidx[k]]
  std::vector<int32_t> Fptr, Fidx;
  std::vector<double> Fw;
  // backward sweep: yF[v] = s[v]*y0[v] + sum_k Bw[k] * yF[Bidx[k]]
  std::vector<int32_t> Bptr, Bidx;
  std::vector<double> Bw;
  std::vector<double> s;

  // level boundaries in position space (nlev + 1 entries)
  std::vector<int32_t> lvl;
  int64_t nlev = 0;
  int64_t tail_level = 0;  // first level of the single-workgroup tail
};

// Convert a caller CSC (int64, base 0/1) to 0-based int32 CSR.
bool csc_to_csr(int64_t n, const int64_t* colptr, const int64_t* rowval,
                const double* nzval, int base, HostCsr& out, std::string& err);

// Build the plan. `tail_width`: levels at the end narrower than this run in
// the single-workgroup tail kernel.
bool build_plan(const HostCsr& L, int64_t nnzF, const int64_t* col,
                const int64_t* colptr, const int64_t* rowval,
                const double* fval, const double* d, int base,
                int64_t tail_width, Plan& P, std::string& err);

}  // namespace lapgpu
