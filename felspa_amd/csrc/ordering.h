// Host-side symbolic work for the GPU ILU: level-of-fill patterns, the
// block-colour ordering, and the stage/block schedule the kernels follow.
#pragma once
#include <cstdint>
#include <vector>

#include "fs_common.h"

namespace fs {

// A schedule partitions the rows (in the factor's order) into contiguous
// blocks; blocks are grouped into stages. Blocks of one stage share no factor
// entry, so a stage runs its blocks in parallel and rows inside a block in
// order. perm[new] = old.
struct Schedule {
  int64_t n = 0;
  std::vector<int32_t> perm, iperm;
  std::vector<int32_t> blk_ptr;    // nb + 1: block b = rows [blk_ptr[b], blk_ptr[b+1])
  std::vector<int32_t> stage_ptr;  // ns + 1: stage s = blocks [stage_ptr[s], stage_ptr[s+1])
  int32_t max_block_rows = 0;
  int n_blocks() const { return (int)blk_ptr.size() - 1; }
  int n_stages() const { return (int)stage_ptr.size() - 1; }
};

// Rows kept in the caller's order, cut into blocks of at most bmax rows,
// one block per stage (stages run one after the other).
Schedule schedule_reference(int64_t n, int bmax);

// Block-colour ordering: rows grouped into blocks (from `hint` block ids or
// by greedy breadth-first aggregation of about `block_rows` rows), blocks
// coloured greedily so that blocks of one colour are more than k+1 graph hops
// apart (so no level-k fill can join them), rows ordered colour by colour,
// block by block, keeping the caller's relative order inside a block.
// Blocks larger than bmax rows are split.
Schedule schedule_block_color(const HostPattern& A, int k, int block_rows, const int32_t* hint,
                              int bmax);

// Merge consecutive blocks of each stage into blocks of up to `rows` rows.
// Only valid when the rows of a stage are mutually independent (point
// colourings, i.e. schedule_block_color with block_rows = 1).
Schedule merge_stage_blocks(const Schedule& s, int rows);

// B = P A P^T for a square pattern (+ values when given), rows sorted;
// map[e] = position in B of entry e of A.
HostCsr permute_square(const HostCsr& A, const std::vector<int32_t>& perm,
                       const std::vector<int32_t>& iperm, std::vector<int64_t>* map);

// Level-of-fill ILU(k) pattern of a square sorted pattern (diagonal added when
// missing), with the reference's rules (include/felspa/linear_algebra/src/
// ilu.implement.h:90-435): k = 0 -> the matrix pattern; k = 1 -> row i gets
// A(i,:) and, for each k < i in A(i,:), the columns of A(k,:) right of k;
// k >= 2 -> generic level recursion, a fill keeping the level it was first
// inserted with. `sched` (may be null) lets independent blocks run in parallel;
// when a row of the schedule would need a row of another block of its own
// stage (the schedule does not separate the factor's blocks), the result is
// empty and *stage_conflict (if given) is set.
HostPattern symbolic_ilu(const HostPattern& A, int k, const Schedule* sched, bool* stage_conflict = nullptr);

// true when no entry of F joins two different blocks of the same stage
bool schedule_is_valid(const HostPattern& F, const Schedule& s);

}  // namespace fs
