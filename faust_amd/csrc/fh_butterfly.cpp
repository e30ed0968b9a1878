// fh_butterfly.cpp -- the butterfly factorization (Faust::butterfly_hierarchical,
// faust_butterfly.hpp:545-860) on the GPU: the schedule of node splits on the
// host, every split one launch of k_butterfly.hip.
//
// Supports are closed-form here.  Support i of L = log2(n) is
// kron(I_{2^i}, kron(ones(2, 2), I_{2^(L-i-1)})) (butterfly_support,
// faust_butterfly.hpp:459-490), and a product of consecutive supports
// i..j-1 is kron(I_{2^i}, kron(ones(2^(j-i)), I_{2^(L-j)})): a node of the
// factorization is the range [a, b) of supports it covers.
//   * 'right' (:545-595): split [i, L) at i+1 for i = 0..L-2 (each solveDTO
//     step; with butterfly supports every equivalence class is a single
//     non-CEC index, so it is the rank-1 lifting);
//   * 'left': split [0, i+1) at i for i = L-1..1;
//   * 'bbtree' (butterfly_hierarchical_balanced, :598-700): the reference's
//     tree of supports, level by level.
#include <algorithm>
#include <memory>
#include <vector>

#include "fh_api_common.hpp"

using namespace fh;

namespace {

struct BNode {
  int a = 0, b = 0;  // supports [a, b)
  DBuf v;            // n x 2^(b-a) values, row-major (the root: the matrix itself)
};
using BNodeP = std::shared_ptr<BNode>;

struct Bfly {
  int n = 0, L = 0;
  BNodeP root;  // [0, L): the matrix to factorize, row-major, columns permuted
  hipStream_t st = nullptr;

  // node [a, b) -> X on [a, c), Y on [c, b)
  std::pair<BNodeP, BNodeP> split(const BNodeP& nd, int c) {
    FH_CHECK(nd->a < c && c < nd->b, "butterfly: split point outside its node");
    auto X = std::make_shared<BNode>();
    auto Y = std::make_shared<BNode>();
    X->a = nd->a;
    X->b = c;
    Y->a = c;
    Y->b = nd->b;
    X->v.alloc((sizeof(double) * (size_t)n) << (c - nd->a));
    Y->v.alloc((sizeof(double) * (size_t)n) << (nd->b - c));
    FH_CHECK(nd->v.p != nullptr, "butterfly: node without values");
    BflyStep s;
    s.in = (const double*)nd->v.p;
    s.in_logB = nd->b - nd->a;
    s.in_logs = L - nd->b;
    s.n = n;
    s.logBx = c - nd->a;
    s.logsx = L - c;
    s.logBy = nd->b - c;
    s.logsy = L - nd->b;
    s.X = (double*)X->v.p;
    s.Y = (double*)Y->v.p;
    bfly_split(s, st);
    return {X, Y};
  }
};

using Range = std::pair<int, int>;

// butterfly_hierarchical_balanced's tree of supports (faust_butterfly.hpp:
// 606-652) as ranges: depth d = ceil(log2 L); level i < d-1 has 2^(i+1)
// nodes, the last level 2 (L - 2^(d-1)) leaves; a node whose children would
// lie past the next level is a leaf itself.
std::vector<std::vector<Range>> bfly_tree(int L) {
  int d = 0;
  while ((1 << d) < L) ++d;
  std::vector<int> sizes;
  for (int i = 0; i + 1 < d; ++i) sizes.push_back(2 << i);
  sizes.push_back(d > 1 ? 2 * (L - sizes[d - 2]) : 2);
  std::vector<std::vector<Range>> tree(d);
  for (int i = d - 1; i >= 0; --i) {
    tree[i].resize(sizes[i]);
    for (int j = 0; j < sizes[i]; ++j) {
      if (i == d - 1) {
        tree[i][j] = {j, j + 1};
      } else if (2 * (j + 1) > sizes[i + 1]) {
        const int k = j + sizes[i + 1] / 2;
        tree[i][j] = {k, k + 1};
      } else {
        const Range &l = tree[i + 1][2 * j], &r = tree[i + 1][2 * j + 1];
        FH_CHECK(l.second == r.first, "butterfly: support tree children not contiguous");
        tree[i][j] = {l.first, r.second};
      }
    }
  }
  FH_CHECK(tree[0][0].first == 0 && tree[0][0].second == tree[0][1].first && tree[0][1].second == L,
           "butterfly: support tree root does not cover the supports");
  return tree;
}

std::vector<BNodeP> bfly_balanced(Bfly& B) {
  const auto tree = bfly_tree(B.L);
  const int d = (int)tree.size();
  auto pr = B.split(B.root, tree[0][0].second);
  B.root.reset();
  std::vector<BNodeP> nxt{pr.first, pr.second}, inp = nxt;
  for (int i = 0; i + 1 < d; ++i) {
    nxt.assign(tree[i + 1].size(), nullptr);
    const int jb = std::min((int)tree[i].size(), (int)tree[i + 1].size() / 2);
    for (int j = 0; j < jb; ++j) {
      FH_CHECK(inp[j]->a == tree[i][j].first && inp[j]->b == tree[i][j].second, "butterfly: tree node mismatch");
      auto ch = B.split(inp[j], tree[i + 1][2 * j].second);
      nxt[2 * j] = ch.first;
      nxt[2 * j + 1] = ch.second;
      if (i < d - 2) inp[j]->v.release();  // only the last level keeps unsplit inputs
    }
    if (i < d - 2) inp = nxt;
  }
  std::vector<BNodeP> facs = nxt;
  if ((int)nxt.size() < B.L)
    for (size_t k = nxt.size() / 2; k < inp.size(); ++k) facs.push_back(inp[k]);
  return facs;
}

std::vector<BNodeP> bfly_directional(Bfly& B, bool right) {
  BNodeP mat = B.root;
  B.root.reset();
  std::vector<BNodeP> facs;
  if (right) {
    for (int i = 0; i + 1 < B.L; ++i) {
      auto pr = B.split(mat, i + 1);
      facs.push_back(pr.first);
      mat = pr.second;
    }
    facs.push_back(mat);
  } else {
    for (int i = B.L - 1; i >= 1; --i) {
      auto pr = B.split(mat, i);
      facs.insert(facs.begin(), pr.second);
      mat = pr.first;
    }
    facs.insert(facs.begin(), mat);
  }
  return facs;
}

// CSR of a leaf [i, i+1) (2 entries per row), built on the device; with
// colmap / colinv (device) its columns are relabelled by colmap and sorted
// within each row
DMatP leaf_csr(const BNode& nd, int n, int L, const int* colmap, const int* colinv, hipStream_t st) {
  auto M = std::make_shared<DMat>();
  bfly_leaf_csr((const double*)nd.v.p, n, L - nd.b, colmap, colinv, *M, st);
  return M;
}

DMatP leaf_dense(const BNode& nd, int n, int L, hipStream_t st) {
  auto M = std::make_shared<DMat>();
  M->kind = K_DENSE;
  M->dtype = FH_F64;
  M->m = n;
  M->n = n;
  M->d.alloc(sizeof(double) * (size_t)n * n);
  FH_HIP(hipMemsetAsync(M->d.p, 0, M->d.bytes, st));
  bfly_scatter_dense((const double*)nd.v.p, n, L - nd.b, (double*)M->d.p, st);
  // the 2n node values land in distinct entries: their nonzeros are the factor's
  M->nnz = count_nnz(nd.v.p, FH_F64, 2 * (int64_t)n, st);
  return M;
}

// the factorization of a device matrix (column-major, or row-major with
// row_major) into a new Faust: Faust::butterfly_hierarchical(A, dir, P,
// mul_perm), faust_butterfly.hpp:804-860
fh_faust_s* butterfly_run(const double* A, size_t lda, int row_major, int n, int dir, const int32_t* perm,
                          int mul_perm, hipStream_t st) {
  FH_CHECK(n >= 4 && (n & (n - 1)) == 0, "The matrix to factorize must be of a size equal to a power of two (>= 4).");
  FH_CHECK(dir == FH_BUTTERFLY_LEFT || dir == FH_BUTTERFLY_RIGHT || dir == FH_BUTTERFLY_BBTREE,
           "dir argument must be 'right', 'left' or 'bbtree'.");
  FH_CHECK(lda >= (size_t)n, "butterfly: bad leading dimension");
  Bfly B;
  B.n = n;
  while ((1 << B.L) < n) ++B.L;
  B.st = st;
  if (dir == FH_BUTTERFLY_BBTREE) {
    // every block of the tree must fit one workgroup: checked before any launch
    const auto tree = bfly_tree(B.L);
    FH_CHECK(bfly_block_fits(tree[0][0].second, B.L - tree[0][0].second),
             "butterfly: bbtree blocks of this order exceed one workgroup's LDS (order <= 32768 supported)");
    for (size_t i = 0; i + 1 < tree.size(); ++i)
      for (size_t j = 0; j < tree[i].size() && 2 * j + 1 < tree[i + 1].size(); ++j) {
        const int a = tree[i][j].first, b = tree[i][j].second, c = tree[i + 1][2 * j].second;
        FH_CHECK(bfly_block_fits(c - a, b - c),
                 "butterfly: bbtree blocks of this order exceed one workgroup's LDS (order <= 32768 supported)");
      }
  }
  // A P^T: column i of A becomes column perm[i], i.e. column c is A's column pinv[c]
  std::vector<int> pinv;
  DBuf pinv_dev, perm_dev;
  if (perm) {
    pinv.assign(n, -1);
    for (int i = 0; i < n; ++i) {
      FH_CHECK(perm[i] >= 0 && perm[i] < n && pinv[perm[i]] < 0, "butterfly: perm is not a permutation of 0..n-1");
      pinv[perm[i]] = i;
    }
    pinv_dev.alloc(sizeof(int) * n);
    FH_HIP(hipMemcpyAsync(pinv_dev.p, pinv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    perm_dev.alloc(sizeof(int) * n);
    FH_HIP(hipMemcpyAsync(perm_dev.p, perm, sizeof(int) * n, hipMemcpyHostToDevice, st));
  }
  B.root = std::make_shared<BNode>();
  B.root->a = 0;
  B.root->b = B.L;
  B.root->v.alloc(sizeof(double) * (size_t)n * n);
  bfly_root(A, lda, row_major, n, (const int*)pinv_dev.p, (double*)B.root->v.p, st);
  std::vector<BNodeP> nodes = dir == FH_BUTTERFLY_BBTREE ? bfly_balanced(B) : bfly_directional(B, dir == FH_BUTTERFLY_RIGHT);
  FH_CHECK((int)nodes.size() == B.L, "butterfly: wrong number of factors");
  for (int k = 0; k < B.L; ++k)
    FH_CHECK(nodes[k]->a == k && nodes[k]->b == k + 1, "butterfly: factor out of order");
  auto r = std::make_unique<fh_faust_s>();
  for (int k = 0; k < B.L; ++k) {
    const bool last = k == B.L - 1;
    if (last && perm && mul_perm) {
      // (F P)[:, c] = F[:, perm[c]]: column j of the last factor moves to pinv[j]
      r->f.push_back(leaf_csr(*nodes[k], n, B.L, (const int*)pinv_dev.p, (const int*)perm_dev.p, st));
    } else if (dir == FH_BUTTERFLY_BBTREE) {
      r->f.push_back(leaf_csr(*nodes[k], n, B.L, nullptr, nullptr, st));
    } else {
      r->f.push_back(leaf_dense(*nodes[k], n, B.L, st));
    }
    nodes[k].reset();
  }
  if (perm && !mul_perm) {
    // P as a factor of its own: row perm[i] holds a 1 at column i
    std::vector<int32_t> rp(n + 1), ci(n);
    std::vector<double> one(n, 1.0);
    for (int row = 0; row < n; ++row) {
      rp[row] = row;
      ci[row] = pinv[row];
    }
    rp[n] = n;
    r->f.push_back(make_csr_from_host(FH_F64, n, n, n, rp.data(), ci.data(), one.data(), st));
  }
  FH_HIP(hipStreamSynchronize(st));
  return r.release();
}

}  // namespace

extern "C" {

int fh_butterfly(const double* A, int n, int dir, const int32_t* perm, int mul_perm, fh_faust_t* out) {
  FH_API_BEGIN
  FH_CHECK(A && out, "null argument");
  FH_CHECK(n >= 4 && (n & (n - 1)) == 0, "The matrix to factorize must be of a size equal to a power of two (>= 4).");
  hipStream_t st = nullptr;
  DBuf Ad(sizeof(double) * (size_t)n * n);
  FH_HIP(hipMemcpyAsync(Ad.p, A, Ad.bytes, hipMemcpyHostToDevice, st));
  *out = butterfly_run((const double*)Ad.p, (size_t)n, 0, n, dir, perm, mul_perm, st);
  FH_API_END
}

int fh_butterfly_device(const double* A, int n, int64_t lda, int row_major, int dir, const int32_t* perm,
                        int mul_perm, fh_faust_t* out, void* stream) {
  FH_API_BEGIN
  FH_CHECK(A && out, "null argument");
  FH_CHECK(lda > 0, "butterfly: bad leading dimension");
  *out = butterfly_run(A, (size_t)lda, row_major, n, dir, perm, mul_perm, (hipStream_t)stream);
  FH_API_END
}

}  // extern "C"
