// Symbolic ILU(k), block-colour ordering and stage schedules (host C++).
#include "ordering.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>
#include <queue>
#include <unordered_map>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace fs {

Schedule schedule_reference(int64_t n, int bmax) {
  Schedule s;
  s.n = n;
  s.perm.resize(n);
  s.iperm.resize(n);
  std::iota(s.perm.begin(), s.perm.end(), 0);
  std::iota(s.iperm.begin(), s.iperm.end(), 0);
  s.blk_ptr.push_back(0);
  for (int64_t r = 0; r < n; r += bmax) s.blk_ptr.push_back((int32_t)std::min<int64_t>(n, r + bmax));
  const int nb = (int)s.blk_ptr.size() - 1;
  for (int b = 0; b <= nb; ++b) s.stage_ptr.push_back(b);
  s.max_block_rows = (int32_t)std::min<int64_t>(n, bmax);
  return s;
}

Schedule schedule_block_color(const HostPattern& A, int k, int block_rows, const int32_t* hint,
                              int bmax) {
  const int64_t n = A.n;
  std::vector<int32_t> blk_of(n, -1);
  std::vector<std::vector<int32_t>> blocks;
  if (hint) {
    std::unordered_map<int32_t, int32_t> id;
    for (int64_t r = 0; r < n; ++r) {
      auto it = id.find(hint[r]);
      int32_t b;
      if (it == id.end()) {
        b = (int32_t)blocks.size();
        id.emplace(hint[r], b);
        blocks.emplace_back();
      } else {
        b = it->second;
      }
      blocks[b].push_back((int32_t)r);
      blk_of[r] = b;
    }
  } else {
    const int target = std::max(1, block_rows);
    std::vector<int32_t> q;
    for (int64_t seed = 0; seed < n; ++seed) {
      if (blk_of[seed] >= 0) continue;
      const int32_t b = (int32_t)blocks.size();
      blocks.emplace_back();
      auto& rows = blocks.back();
      q.clear();
      q.push_back((int32_t)seed);
      blk_of[seed] = b;
      rows.push_back((int32_t)seed);
      for (size_t h = 0; h < q.size() && (int)rows.size() < target; ++h) {
        const int32_t u = q[h];
        for (int64_t p = A.rp[u]; p < A.rp[u + 1] && (int)rows.size() < target; ++p) {
          const int32_t v = A.ci[p];
          if (blk_of[v] >= 0) continue;
          blk_of[v] = b;
          rows.push_back(v);
          q.push_back(v);
        }
      }
      std::sort(rows.begin(), rows.end());
    }
  }
  // split blocks larger than bmax
  {
    std::vector<std::vector<int32_t>> split;
    for (auto& rows : blocks) {
      for (size_t o = 0; o < rows.size(); o += bmax)
        split.emplace_back(rows.begin() + o, rows.begin() + std::min(rows.size(), o + (size_t)bmax));
    }
    blocks.swap(split);
    // blocks ordered by their first row
    std::sort(blocks.begin(), blocks.end(),
              [](const std::vector<int32_t>& a, const std::vector<int32_t>& b) { return a[0] < b[0]; });
    for (size_t b = 0; b < blocks.size(); ++b)
      for (int32_t r : blocks[b]) blk_of[r] = (int32_t)b;
  }
  const int64_t nb = (int64_t)blocks.size();
  // blocks reachable within k+1 hops
  std::vector<std::vector<int32_t>> adj(nb);
#pragma omp parallel
  {
    std::vector<int32_t> stamp(n, -1);
    std::vector<int32_t> front, next, nbs;
#pragma omp for schedule(dynamic, 16)
    for (int64_t b = 0; b < nb; ++b) {
      front.assign(blocks[b].begin(), blocks[b].end());
      for (int32_t r : front) stamp[r] = (int32_t)b;
      nbs.clear();
      for (int layer = 0; layer <= k; ++layer) {
        next.clear();
        for (int32_t u : front)
          for (int64_t p = A.rp[u]; p < A.rp[u + 1]; ++p) {
            const int32_t v = A.ci[p];
            if (stamp[v] == (int32_t)b) continue;
            stamp[v] = (int32_t)b;
            next.push_back(v);
            nbs.push_back(blk_of[v]);
          }
        front.swap(next);
      }
      std::sort(nbs.begin(), nbs.end());
      nbs.erase(std::unique(nbs.begin(), nbs.end()), nbs.end());
      adj[b] = nbs;
    }
  }
  // a nonsymmetric pattern reaches a block from blocks it does not reach
  // itself: join every list with its transpose, so two blocks within k+1
  // hops of each other in either direction never share a colour
  {
    std::vector<int64_t> tc(nb + 1, 0);
    for (int64_t b = 0; b < nb; ++b)
      for (int32_t a : adj[b]) tc[a + 1]++;
    for (int64_t b = 0; b < nb; ++b) tc[b + 1] += tc[b];
    std::vector<int32_t> tl(tc[nb]);
    std::vector<int64_t> fill(tc.begin(), tc.end() - 1);
    for (int64_t b = 0; b < nb; ++b)
      for (int32_t a : adj[b]) tl[fill[a]++] = (int32_t)b;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < nb; ++b) {
      auto& l = adj[b];
      const size_t own = l.size();
      l.insert(l.end(), tl.begin() + tc[b], tl.begin() + tc[b + 1]);
      std::inplace_merge(l.begin(), l.begin() + own, l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
    }
  }
  // greedy colouring in block order
  std::vector<int32_t> color(nb, -1);
  int ncolor = 0;
  std::vector<int64_t> used;
  for (int64_t b = 0; b < nb; ++b) {
    used.assign(ncolor + 1, -1);
    for (int32_t a : adj[b])
      if (color[a] >= 0) used[color[a]] = b;
    int c = 0;
    while (c < ncolor && used[c] == b) ++c;
    color[b] = c;
    ncolor = std::max(ncolor, c + 1);
  }
  Schedule s;
  s.n = n;
  s.perm.reserve(n);
  s.blk_ptr.push_back(0);
  s.stage_ptr.push_back(0);
  std::vector<std::vector<int32_t>> by_color(ncolor);
  for (int64_t b = 0; b < nb; ++b) by_color[color[b]].push_back((int32_t)b);
  for (int c = 0; c < ncolor; ++c) {
    for (int32_t b : by_color[c]) {
      for (int32_t r : blocks[b]) s.perm.push_back(r);
      s.blk_ptr.push_back((int32_t)s.perm.size());
      s.max_block_rows = std::max<int32_t>(s.max_block_rows, (int32_t)blocks[b].size());
    }
    s.stage_ptr.push_back((int32_t)s.blk_ptr.size() - 1);
  }
  s.iperm.resize(n);
  for (int64_t i = 0; i < n; ++i) s.iperm[s.perm[i]] = (int32_t)i;
  return s;
}

Schedule merge_stage_blocks(const Schedule& s, int rows) {
  Schedule m = s;
  m.blk_ptr.assign(1, 0);
  m.stage_ptr.assign(1, 0);
  m.max_block_rows = 0;
  for (int st = 0; st < s.n_stages(); ++st) {
    const int32_t a = s.blk_ptr[s.stage_ptr[st]], b = s.blk_ptr[s.stage_ptr[st + 1]];
    for (int32_t r = a; r < b; r += rows) {
      const int32_t e = std::min<int32_t>(b, r + rows);
      m.blk_ptr.push_back(e);
      m.max_block_rows = std::max(m.max_block_rows, e - r);
    }
    m.stage_ptr.push_back((int32_t)m.blk_ptr.size() - 1);
  }
  return m;
}

HostCsr permute_square(const HostCsr& A, const std::vector<int32_t>& perm,
                       const std::vector<int32_t>& iperm, std::vector<int64_t>* map) {
  const int64_t n = A.n_rows;
  HostCsr B;
  B.n_rows = B.n_cols = n;
  B.rp.assign(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) B.rp[r + 1] = B.rp[r] + (A.rp[perm[r] + 1] - A.rp[perm[r]]);
  B.ci.resize(B.rp[n]);
  const bool vals = !A.v.empty();
  if (vals) B.v.resize(B.rp[n]);
  if (map) map->assign(A.nnz(), -1);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, int64_t>> tmp;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
      const int64_t o = perm[r];
      tmp.clear();
      for (int64_t p = A.rp[o]; p < A.rp[o + 1]; ++p) tmp.emplace_back(iperm[A.ci[p]], p);
      std::sort(tmp.begin(), tmp.end());
      int64_t q = B.rp[r];
      for (auto& e : tmp) {
        B.ci[q] = e.first;
        if (vals) B.v[q] = A.v[e.second];
        if (map) (*map)[e.second] = q;
        ++q;
      }
    }
  }
  return B;
}

namespace {

using RowList = std::vector<std::vector<int32_t>>;

// row i of A with the diagonal inserted when missing (sorted)
void row_with_diag(const HostPattern& A, int64_t i, std::vector<int32_t>& out) {
  out.assign(A.ci.begin() + A.rp[i], A.ci.begin() + A.rp[i + 1]);
  if (!std::binary_search(out.begin(), out.end(), (int32_t)i))
    out.insert(std::lower_bound(out.begin(), out.end(), (int32_t)i), (int32_t)i);
}

}  // namespace

HostPattern symbolic_ilu(const HostPattern& A, int k, const Schedule* sched, bool* stage_conflict) {
  const int64_t n = A.n;
  if (stage_conflict) *stage_conflict = false;
  RowList rows(n);
  if (k <= 1) {
#pragma omp parallel
    {
      std::vector<int32_t> acc, ri;
#pragma omp for schedule(dynamic, 512)
      for (int64_t i = 0; i < n; ++i) {
        row_with_diag(A, i, ri);
        acc = ri;
        if (k == 1)
          for (int32_t kk : ri) {
            if (kk >= i) break;
            for (int64_t p = A.rp[kk]; p < A.rp[kk + 1]; ++p)
              if (A.ci[p] > kk) acc.push_back(A.ci[p]);
          }
        std::sort(acc.begin(), acc.end());
        acc.erase(std::unique(acc.begin(), acc.end()), acc.end());
        rows[i].swap(acc);
      }
    }
  } else {
    // Row i is stored as [L part by column][diagonal][U part by (level,
    // column)] with the levels alongside, and ucnt[i k + l] = number of its U
    // entries of level <= l (l < k). A left neighbour kk of level lk can only
    // insert columns of row kk whose level is <= k - lk - 1 (any other gives a
    // level above k), so only that prefix of row kk's U part is scanned. The
    // insertions of one kk are independent of the order in which its columns
    // are visited, so the pattern and levels equal those of a full scan.
    std::vector<std::vector<uint8_t>> lev(n);
    std::vector<int32_t> dpos(n);
    const uint8_t INV = 255;
    if (k > 254) k = 254;
    std::vector<uint32_t> ucnt((size_t)n * (size_t)k);
    // stage-parallel rows may only read rows of earlier stages or of their
    // own block; a schedule that breaks this is reported, never read through
    // per row: the first row of its block and of its stage (contiguous
    // ranges), so a left neighbour kk < i is in another block of the same
    // stage iff stage start <= kk < block start
    std::vector<int32_t> blk_first, stage_first;
    const bool staged = sched && sched->n == n;
    if (staged) {
      blk_first.resize(n);
      stage_first.resize(n);
      for (int st = 0; st < sched->n_stages(); ++st) {
        const int32_t s0 = sched->blk_ptr[sched->stage_ptr[st]];
        for (int b = sched->stage_ptr[st]; b < sched->stage_ptr[st + 1]; ++b)
          for (int32_t r = sched->blk_ptr[b]; r < sched->blk_ptr[b + 1]; ++r) {
            blk_first[r] = sched->blk_ptr[b];
            stage_first[r] = s0;
          }
      }
    }
    int conflict = 0;
    auto do_row = [&](int64_t i, std::vector<uint8_t>& mark, std::vector<int32_t>& cols, std::vector<int32_t>& ri,
                      std::vector<uint64_t>& up, int& bad) {
      row_with_diag(A, i, ri);
      cols.clear();
      std::priority_queue<int32_t, std::vector<int32_t>, std::greater<int32_t>> heap;
      for (int32_t c : ri) {
        mark[c] = 0;
        cols.push_back(c);
        if (c < i) heap.push(c);
      }
      while (!heap.empty()) {
        const int32_t kk = heap.top();
        heap.pop();
        const int lk = mark[kk];
        if (lk >= k) continue;
        if (staged && kk >= stage_first[i] && kk < blk_first[i]) {
          bad = 1;  // row kk may still be in the making on another thread
          continue;
        }
        const auto& rk = rows[kk];
        const auto& lv = lev[kk];
        const size_t p0 = (size_t)dpos[kk] + 1;
        const size_t p1 = p0 + ucnt[(size_t)kk * k + (size_t)(k - lk - 1)];
        for (size_t p = p0; p < p1; ++p) {
          const int32_t j = rk[p];
          if (mark[j] != INV) continue;  // an existing entry keeps its first level
          const int L = lv[p] + lk + 1;
          if (L <= k) {
            mark[j] = (uint8_t)L;
            cols.push_back(j);
            if (j < i) heap.push(j);
          }
        }
      }
      // L part by column, the diagonal, U part by (level, column)
      up.clear();
      std::vector<int32_t> row;
      row.reserve(cols.size());
      for (int32_t c : cols) {
        if (c < i) row.push_back(c);
        else if (c > i) up.push_back(((uint64_t)mark[c] << 32) | (uint32_t)c);
      }
      std::sort(row.begin(), row.end());
      std::sort(up.begin(), up.end());
      const size_t nl = row.size();
      row.push_back((int32_t)i);
      for (uint64_t e : up) row.push_back((int32_t)(uint32_t)e);
      std::vector<uint8_t> lvl(row.size());
      uint32_t* uc = ucnt.data() + (size_t)i * k;
      for (int l = 0; l < k; ++l) uc[l] = 0;
      for (size_t p = 0; p < row.size(); ++p) {
        const int32_t c = row[p];
        lvl[p] = mark[c];
        mark[c] = INV;
        if (c > i)
          for (int l = lvl[p]; l < k; ++l) uc[l]++;
      }
      dpos[i] = (int32_t)nl;
      rows[i].swap(row);
      lev[i].swap(lvl);
    };
    if (staged) {
      for (int s = 0; s < sched->n_stages() && !conflict; ++s) {
        const int b0 = sched->stage_ptr[s], b1 = sched->stage_ptr[s + 1];
#pragma omp parallel reduction(| : conflict)
        {
          std::vector<uint8_t> mark(n, INV);
          std::vector<int32_t> cols, ri;
          std::vector<uint64_t> up;
          int bad = 0;
#pragma omp for schedule(dynamic, 1)
          for (int b = b0; b < b1; ++b)
            for (int32_t i = sched->blk_ptr[b]; i < sched->blk_ptr[b + 1]; ++i) do_row(i, mark, cols, ri, up, bad);
          conflict |= bad;
        }
      }
    } else {
      std::vector<uint8_t> mark(n, INV);
      std::vector<int32_t> cols, ri;
      std::vector<uint64_t> up;
      int bad = 0;
      for (int64_t i = 0; i < n; ++i) do_row(i, mark, cols, ri, up, bad);
    }
    if (conflict) {
      if (stage_conflict) *stage_conflict = true;
      return HostPattern{};
    }
    // U parts back to column order
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) std::sort(rows[i].begin() + dpos[i] + 1, rows[i].end());
  }
  HostPattern F;
  F.n = n;
  F.rp.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) F.rp[i + 1] = F.rp[i] + (int64_t)rows[i].size();
  F.ci.resize(F.rp[n]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) std::copy(rows[i].begin(), rows[i].end(), F.ci.begin() + F.rp[i]);
  return F;
}

bool schedule_is_valid(const HostPattern& F, const Schedule& s) {
  // stages and blocks are contiguous row ranges: an entry (i, j) joins two
  // blocks of one stage iff j lies in i's stage range but not in its block
  bool ok = true;
#pragma omp parallel for schedule(dynamic, 1) reduction(&& : ok)
  for (int st = 0; st < s.n_stages(); ++st) {
    const int32_t s0 = s.blk_ptr[s.stage_ptr[st]], s1 = s.blk_ptr[s.stage_ptr[st + 1]];
    for (int b = s.stage_ptr[st]; b < s.stage_ptr[st + 1]; ++b) {
      const int32_t b0 = s.blk_ptr[b], b1 = s.blk_ptr[b + 1];
      for (int32_t i = b0; i < b1; ++i)
        for (int64_t p = F.rp[i]; p < F.rp[i + 1]; ++p) {
          const int32_t j = F.ci[p];
          if (j >= s0 && j < s1 && (j < b0 || j >= b1)) ok = false;
        }
    }
  }
  return ok;
}

}  // namespace fs
