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

// One entry of a slot program: w * y[src].  A lane of a step sums up to
// E = 2^elog entries (the step's "planes"), then the lanes of an aligned
// segment of 2^seglog lanes are summed and the segment's last lane adds the
// sum to y[row] (lead = 1: plain add; lead = 2: atomic add, for rows split
// in chunks and summed in any order) or stores it to the partial slot
// y[row] (lead = 3: a chunk of a row that a later fix-up step sums in chunk
// order).  meta packs row (bits 0-25), seglog (bits 26-28) and lead (bits
// 29-31); it is read from plane 0 only.  Indices are local to the program
// (position - T0); partial slots follow them, [NT, NT + npart).
struct TailSlot {
  double w;
  uint32_t src;
  uint32_t meta;
};
constexpr uint32_t kRowMask = (1u << 26) - 1;
inline uint32_t slot_meta(uint32_t row, uint32_t seglog, uint32_t lead) {
  return row | (seglog << 26) | (lead << 29);
}
static_assert(sizeof(TailSlot) == 16, "TailSlot is 16 bytes");

// A step = up to kTailLanes lanes run by the tail workgroup at once.  Its
// slots are E planes of `count` slots each: lane t's entry e is slot
// off + e * count + t.  meta = count (active lanes, a multiple of 64)
// | end << 16 (the last step of a level: a workgroup barrier follows)
// | elog << 17 (E = 2^elog, at most kTailMaxElog)
// | maxlog << 19 (the largest seglog in the step).
struct TailStep {
  int32_t off;   // first slot
  int32_t meta;
};
constexpr int kTailMaxElog = 2;
// Partial slots a program may use (past its own indices; the kernels' y
// arrays hold them).
constexpr int64_t kPartCap = 1024;
inline int32_t step_meta(int64_t count, bool end, int elog, int maxlog) {
  return (int32_t)(count | ((end ? 1 : 0) << 16) | (elog << 17) | (maxlog << 19));
}

// Lanes of one step of the tail's slot programs (the tail workgroup's
// threads).  The whole-graph programs of the batched solver use their own
// workgroup width.
#ifndef LAPGPU_TAIL_THREADS
#define LAPGPU_TAIL_THREADS 512
#endif
constexpr int kTailLanes = LAPGPU_TAIL_THREADS;
// Lanes per step the extent rule of build_tail counts steps in.
constexpr int kExtentLanes = 1024;
#ifndef LAPGPU_TAIL_CAP
#define LAPGPU_TAIL_CAP 18432
#endif
constexpr int64_t kTailCap = LAPGPU_TAIL_CAP;  // tail vertices held in LDS (147 KiB)

// The tail [T0, n) is run by one workgroup.  With a dense block, its last
// n - T1 vertices (T1 = T0 + NT1, a level boundary) are applied as one
// matrix instead (DenseBlock): fsteps then run the levels of [T0, T1) and one
// more level that gathers the block's inputs from [T0, T1), and bsteps the
// levels of [T0, T1), whose rows may read the block's results.
struct TailProgram {
  int64_t T0 = 0, NT = 0;
  int64_t NT1 = 0;  // vertices before the dense block (NT without one)
  int elog = 0;     // every step of both sweeps has 2^elog entries per lane
  int64_t npart = 0;  // partial slots past [0, NT) (at most kPartCap)
  std::vector<TailStep> fsteps, bsteps;
  std::vector<TailSlot> fslots, bslots;
  // forward in-edges of tail rows from outside the tail (CSR over tail rows)
  std::vector<int32_t> Eptr, Eidx;
  std::vector<double> Ew;
};

// The dense block [T1, n) of the tail: its map from the forward inputs u
// (r plus every forward in-edge from before T1) to yF is the symmetric
// matrix D = (I - W^T)^-1 S (I - W)^-1 of the block's own levels, stored
// less its constant part s_root 1 1^T (which only shifts yF by a constant
// that the mean removal cancels).  Its programs (indices local to T1) are
// run once per unit vector to build D.
struct DenseBlock {
  int64_t T1 = 0, K = 0;  // K = 0: no dense block
  int64_t level = 0;      // first level of the block
  int elog = 0;
  int64_t npart = 0;  // partial slots past [0, K)
  std::vector<TailStep> fsteps, bsteps;
  std::vector<TailSlot> fslots, bslots;
};

// Choice of the dense block (build_tail): the split level minimising
// (steps left to the slot programs) x step_us + 8 K^2 bytes / gbps (+ two
// launches when a sparse part remains), with K <= max_k (0: never dense).
struct DenseModel {
  int64_t max_k = 0;
  // D u adds terms of size s_max |u| that cancel down to the result, while
  // the sweeps difference first and scale after: a block whose diagonal
  // scales s span more than max_spread (a rounding error of about eps x
  // spread) is left to the sweeps
  double max_spread = 1e12;
  double step_us = 0.45;   // one dependent 512-lane step of the tail (FP64, measured)
  double gbps = 5000.0;    // dense matrix-vector product streaming rate
  double launch_us = 3.0;  // one more dependent launch
};

// Groups of consecutive wide levels computed by one launch each.  Row v of
// a group gets a composite row over the vertices outside the group plus
// "self" terms of the group's rows, with indices idx < n naming A and
// idx >= n naming B[idx - n]:
//   forward:  y0[v] = sum w * (idx < n ? y0[idx] : r[idx - n])
//   backward: yF[v] = sum w * (idx < n ? yF[idx] : y0[idx - n])
// (a one-level group holds exactly the original gathers plus the r[v] or
// s[v]*y0[v] term).
struct LevelBlocks {
  std::vector<int32_t> lev;  // group g = levels [lev[g], lev[g+1])
  std::vector<int32_t> fptr, fidx, bptr, bidx;  // CSR over positions [0, T0)
  std::vector<double> fw, bw;
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
  TailProgram tail;
  DenseBlock dense;
  LevelBlocks blocks;
};

// Convert a caller CSC (int64, base 0/1) to 0-based int32 CSR.
bool csc_to_csr(int64_t n, const int64_t* colptr, const int64_t* rowval,
                const double* nzval, int base, HostCsr& out, std::string& err);

// Build the plan. The tail is the longest suffix of levels that fits in LDS
// (kTailCap vertices) and whose levels need at most `max_tail_steps` steps.
bool build_plan(const HostCsr& L, int64_t nnzF, const int64_t* col,
                const int64_t* colptr, const int64_t* rowval,
                const double* fval, const double* d, int base,
                int64_t max_tail_steps, Plan& P, std::string& err,
                const DenseModel& dense = DenseModel());

// (Re)build P.tail as the longest suffix of levels with at most `cap`
// vertices whose levels need at most `max_steps` steps each, and P.dense by
// `dense`.  With both limits at INT64_MAX the program covers the whole sweep
// (tail_level 0).  Entries per lane are chosen for the whole program, at
// most 2^max_elog.
// With det (the default) every row sum of the programs is formed in a
// fixed order (partial slots and fix-up steps for rows split in chunks), so
// that the sweeps give the same bits on every run; without, split rows add
// their chunks atomically.
bool build_tail(Plan& P, int64_t max_steps, int64_t cap, int max_elog = kTailMaxElog,
                int lanes = kTailLanes, const DenseModel& dense = DenseModel(), bool det = true);

// The plan file (lapgpu_planio.cpp): the whole plan, including the level
// groups, with a checksum; load_plan refuses a file that fails it or whose
// shapes or indices do not fit together.
bool save_plan(const Plan& P, const std::string& path, std::string& err);
bool load_plan(const std::string& path, Plan& P, std::string& err);

// Group the levels before the tail into launches of up to `max_depth`
// levels while a group's composite rows stay within `growth` times the
// entries of its levels (and `max_extra` extra entries).
void build_blocks(Plan& P, int max_depth, double growth, int64_t max_extra);

}  // namespace lapgpu
