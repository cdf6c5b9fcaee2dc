// simba_host.h -- host-side pieces of the SIMBA search shared by the
// per-instance and batched drivers.
#pragma once
#include <vector>

#include "simba_kernels.h"

namespace mba {

inline uint64_t wmask(int W) { return W == 64 ? ~0ull : 0xffffffffull; }

// The sample whose values reject the most candidates (see simba_host.cpp);
// `skip` excludes one sample (-1: none).  Returns -1 when none is left.
int pick_filter_sample(const uint64_t *lv, const uint64_t *outv, int n, int L, int W,
                       int skip = -1);
int choose_kt(const CountTable &T, int max_size, int W, uint64_t budget);
// A binary (op, left size) block of size s and the pair launch covering it.
struct BlockRef {
  int op, split;        // root operator, left child size
  uint64_t begin, end;  // rank range within the size
  int pair_j;           // the pair {pair_j, s-1-pair_j}
  int inst;             // operator instance within that pair
};
// Binary blocks of size s in rank order (nocache.cu:52-79).
std::vector<BlockRef> binary_blocks(const CountTable &T, int s);
// The splits {j, s-1-j} of size s in the order learned block scores
// (scores[op * 32 + left size]) propose: by the probability their blocks
// get (softmax over the size's blocks) per pair checked, highest first.
std::vector<int> learned_split_order(const CountTable &T, int s, const float *scores);
// Below this many trees a size is swept with fused pairs and one sync.
constexpr uint64_t kBlockModeMin = 1ull << 31;
// The same threshold for the learned order, whose point is to stop early.
constexpr uint64_t kLearnedBlockMin = 1ull << 26;

// The B-side trees of size m > Kt in rank order as blocks formed from the
// tables V[1..Kt] (table references resolved through the instance's
// TableView; see b_blocks in simba_host.cpp), for tiles of 256 * rb B
// values; block_tiles: the tiles of one block.
void b_blocks(const CountTable &T, int Kt, int m, uint64_t start, uint64_t stride,
              const std::vector<BStep> &steps, int rb, std::vector<BBlock> &out);
uint64_t block_tiles(const BBlock &b, int rb);

// Windows for the checks of base size sb (chains up to max_size).
CheckTargets targets_for(uint64_t t0, int sb, int max_size, int W);
// Left child size of the root of an RPN (0 for unary roots and leaves).
int root_split(const uint8_t *rpn, int size);
// Fills the tree fields of a result from (size, rank).
void fill_result(const CountTable &T, int size, uint64_t rank, int n_vars,
                 const uint64_t *consts, int W, mba_result &res);

}  // namespace mba
