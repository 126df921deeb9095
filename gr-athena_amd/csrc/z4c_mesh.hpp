// z4c_mesh.hpp -- host-side plan of the ghost update across refinement levels.
//
// The blocks of a context (equal cell counts, spacings h0 * 2^-L) are placed
// on a block tree of the domain [lo, hi] (periodic or open per axis); for
// every block the 26 neighbour directions are resolved to same-level,
// coarser or finer blocks (mesh_topology::SearchAndSetNeighbors,
// src/mesh/mesh_topology.cpp:278-960: faces, then x1x2, x1x3, x2x3 edges,
// then corners; a coarser edge or corner neighbour only where the block sits
// in that corner of its parent).  The ghost update of one stage
// (SEND_Z4C / RECV_Z4C / SETB_Z4C / PROLONG of the GR_Z4c task list) then
// becomes a fixed linear map, which the plan stores as
//
//  * one gather list per destination vertex (fine ghosts, coarse-buffer
//    ghosts, and interior boundary vertices whose node multiplicity exceeds
//    one): the vertex's own value if it is interior (ghosts start from zero,
//    NodeMultiplicity::ZeroGhosts), then the contributions in the order
//    CommChannel::Unpack adds them (src/comm/comm_channel.cpp:1277-1472;
//    an overwrite from a coarser neighbour restarts the list), and the
//    factor 1/multiplicity (NodeMultiplicity::ApplyDivision, one table for
//    the fine array and the coarse buffer, src/comm/node_multiplicity.cpp);
//  * the prolongation boxes (ProlongationRange of every coarser neighbour,
//    src/comm/refinement_ops.cpp:32-176).
//
// Every source of a gather list is an interior vertex: of a fine array
// (same-level slabs, LoadToFiner half-slabs) or of a coarse buffer's
// restricted interior (coarse payloads, LoadToCoarser slabs).  Index ranges
// are those of src/comm/index_utilities.{hpp,cpp} for Sampling::VC with
// NGHOST = NCGHOST = 4.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

namespace z4c_mesh {

constexpr int MNG = 4;   // fine ghosts
constexpr int MCNG = 4;  // coarse ghosts

struct Geo {
  int nx[3];         // cells per axis
  double dx[3];      // spacing
  double xmin[3];    // first interior vertex
};

struct Nbr {
  int ox[3];
  int fi1, fi2;
  int rel;  // -1 coarser, 0 same level, +1 finer
  int id;
};

// One axis of a block's vertex-centred index set (fine or coarse).
struct Axis {
  int vs, ve, ms, me, ps, pe, gs, ge;
};

inline Axis axis_of(int cells, int ng) {
  Axis a;
  a.vs = ng;
  a.ve = ng + cells;
  a.ms = 0;
  a.me = ng - 1;
  a.ps = cells + ng + 1;
  a.pe = cells + 2 * ng;
  a.gs = a.vs + ng;
  a.ge = a.ve - ng;
  return a;
}

struct Range {
  int s, e;
};

// --- per-axis ranges of src/comm/index_utilities.hpp (VC) ------------------
inline Range load_same(int ox, const Axis& a) { return {ox > 0 ? a.ge : a.vs, ox < 0 ? a.gs : a.ve}; }
inline Range set_same(int ox, const Axis& a) {
  if (ox == 0) return {a.vs, a.ve};
  return ox > 0 ? Range{a.ve, a.pe} : Range{a.ms, a.vs};
}
inline Range load_to_coarser(int ox, const Axis& c) {
  if (ox > 0) return {c.ve - MNG, c.ve};
  if (ox < 0) return {c.vs, c.vs + MNG};
  return {c.vs, c.ve};
}
inline Range set_from_finer(int ox, const Axis& a, int fi, int half) {
  if (ox == 0) return fi == 1 ? Range{a.vs + half, a.ve} : Range{a.vs, a.ve - half};
  return ox > 0 ? Range{a.ve, a.pe} : Range{a.ms, a.vs};
}
inline Range load_to_finer(int ox, const Axis& a, int fi, int half) {
  if (ox > 0) return {a.ve - MCNG, a.ve - 1};
  if (ox < 0) return {a.vs + 1, a.vs + MCNG};
  return fi == 1 ? Range{a.vs + half, a.ve} : Range{a.vs, a.ve - half};
}
inline Range set_from_coarser(int ox, const Axis& c, long long lx) {
  if (ox == 0) return (lx & 1) == 0 ? Range{c.vs, c.pe} : Range{c.ms, c.ve};
  return ox > 0 ? Range{c.ps, c.pe} : Range{c.ms, c.me};
}
// fi of axis d (ox == 0 there): x1 always fi1, x2 fi1 when ox1 != 0, x3 fi1
// when ox1 and ox2 != 0 (index_utilities.cpp:769-772)
inline int fi_of(const Nbr& n, int d) {
  const bool f1 = d == 0 || (d == 1 && n.ox[0] != 0) || (d == 2 && n.ox[0] != 0 && n.ox[1] != 0);
  return f1 ? n.fi1 : n.fi2;
}

// --- node multiplicity in the 7-zone compact space ---------------------------
inline Range compact(int rel, int ox, long long lx, int fi) {
  if (rel == 0) return ox == 0 ? Range{1, 5} : (ox > 0 ? Range{5, 6} : Range{0, 1});
  if (rel < 0) {
    if (ox == 0) return (lx & 1) == 0 ? Range{1, 6} : Range{0, 5};
    return ox > 0 ? Range{6, 6} : Range{0, 0};
  }
  if (ox == 0) return fi == 1 ? Range{3, 5} : Range{1, 3};
  return ox > 0 ? Range{5, 6} : Range{0, 1};
}
// compact zone of index i on an axis (pivots ms, vs, vs+1, vs+half, ve-1, ve, pe)
inline int zone_of(int i, const Axis& a) {
  const int half = (a.ve - a.vs) / 2;
  if (i < a.vs) return 0;
  if (i == a.vs) return 1;
  if (i < a.vs + half) return 2;
  if (i == a.vs + half) return 3;
  if (i < a.ve) return 4;
  if (i == a.ve) return 5;
  return 6;
}

// --- the block tree ------------------------------------------------------------
struct Tree {
  int n = 0;
  int nx[3] = {0, 0, 0};
  bool periodic[3] = {true, true, true};
  long long nrb0[3] = {0, 0, 0};  // blocks per axis at level 0
  std::vector<int> level;
  std::vector<std::array<long long, 3>> loc;
  std::unordered_map<unsigned long long, int> leaf;
  std::vector<std::vector<Nbr>> nbrs;
  std::vector<char> all_same;

  static unsigned long long key(int L, long long a, long long b, long long c) {
    return ((unsigned long long)L << 60) | ((unsigned long long)a << 40) |
           ((unsigned long long)b << 20) | (unsigned long long)c;
  }
  int find(int L, long long a, long long b, long long c) const {
    auto it = leaf.find(key(L, a, b, c));
    return it == leaf.end() ? -1 : it->second;
  }
  // wraps t at level L; false beyond an open face
  bool wrap(int L, long long t[3]) const {
    for (int d = 0; d < 3; ++d) {
      const long long m = nrb0[d] << L;
      if (t[d] < 0 || t[d] >= m) {
        if (!periodic[d]) return false;
        t[d] = ((t[d] % m) + m) % m;
      }
    }
    return true;
  }
  // 0 same-level leaf, +1 refined, -1 covered by the parent; 2 unbalanced
  int cover(int L, const long long t[3]) const {
    if (find(L, t[0], t[1], t[2]) >= 0) return 0;
    for (int c = 0; c < 8; ++c)
      if (find(L + 1, 2 * t[0] + (c & 1), 2 * t[1] + ((c >> 1) & 1), 2 * t[2] + (c >> 2)) >= 0)
        return 1;
    if (L > 0 && find(L - 1, t[0] >> 1, t[1] >> 1, t[2] >> 1) >= 0) return -1;
    return 2;
  }

  std::string build(const std::vector<Geo>& g, const double lo[3], const double hi[3],
                    const int per[3]) {
    n = (int)g.size();
    for (int d = 0; d < 3; ++d) {
      nx[d] = g[0].nx[d];
      periodic[d] = per[d] != 0;
    }
    double hmax = 0.0;
    for (const Geo& b : g) {
      for (int d = 0; d < 3; ++d)
        if (b.nx[d] != nx[d]) return "all blocks need the same number of cells per axis";
      hmax = std::max(hmax, b.dx[0]);
    }
    for (int d = 0; d < 3; ++d) {
      if (nx[d] % 2 != 0 || nx[d] < 2 * MCNG)
        return "blocks need an even number of at least 8 cells per axis";
      const double r = (hi[d] - lo[d]) / (nx[d] * hmax);
      nrb0[d] = std::llround(r);
      if (nrb0[d] <= 0 || std::fabs(r - (double)nrb0[d]) > 1e-6)
        return "the domain is not a whole number of blocks of the coarsest spacing";
    }
    level.resize(n);
    loc.resize(n);
    for (int b = 0; b < n; ++b) {
      const double r = std::log2(hmax / g[b].dx[0]);
      const int L = (int)std::lround(r);
      if (std::fabs(r - L) > 1e-9 || L > 15) return "spacing is not the coarsest halved L times";
      for (int d = 0; d < 3; ++d) {
        if (std::fabs(g[b].dx[d] - std::ldexp(hmax, -L)) > 1e-12 * hmax)
          return "spacing differs between axes or from the level's";
        const double q = (g[b].xmin[d] - lo[d]) / (nx[d] * g[b].dx[d]);
        loc[b][d] = std::llround(q);
        if (std::fabs(q - (double)loc[b][d]) > 1e-6 || loc[b][d] < 0 ||
            loc[b][d] >= (nrb0[d] << L))
          return "block " + std::to_string(b) + " does not sit on its level's block lattice";
      }
      level[b] = L;
      if (!leaf.emplace(key(L, loc[b][0], loc[b][1], loc[b][2]), b).second)
        return "two blocks at the same place";
    }
    nbrs.assign(n, {});
    all_same.assign(n, 1);
    for (int b = 0; b < n; ++b) {
      std::string e = search(b);
      if (!e.empty()) return e;
      for (const Nbr& q : nbrs[b])
        if (q.rel != 0) all_same[b] = 0;
    }
    return "";
  }

  std::string search(int b) {
    const int L = level[b];
    const long long* lx = loc[b].data();
    int my[3];
    for (int d = 0; d < 3; ++d) my[d] = 2 * (int)(lx[d] & 1) - 1;
    std::vector<Nbr>& out = nbrs[b];
    auto target = [&](const int ox[3], long long t[3]) {
      for (int d = 0; d < 3; ++d) t[d] = lx[d] + ox[d];
      return wrap(L, t);
    };
    auto child = [&](const long long t[3], int c0, int c1, int c2) {
      return find(L + 1, 2 * t[0] + c0, 2 * t[1] + c1, 2 * t[2] + c2);
    };
    auto push = [&](const int ox[3], int rel, int id, int fi1, int fi2) {
      Nbr q;
      for (int d = 0; d < 3; ++d) q.ox[d] = ox[d];
      q.rel = rel;
      q.id = id;
      q.fi1 = fi1;
      q.fi2 = fi2;
      out.push_back(q);
    };
    auto same_or_coarser = [&](const long long t[3], int kind) {
      return kind == 0 ? find(L, t[0], t[1], t[2]) : find(L - 1, t[0] >> 1, t[1] >> 1, t[2] >> 1);
    };
    // faces
    for (int ax = 0; ax < 3; ++ax)
      for (int s = -1; s <= 1; s += 2) {
        int ox[3] = {0, 0, 0};
        ox[ax] = s;
        long long t[3];
        if (!target(ox, t)) continue;
        const int kind = cover(L, t);
        if (kind == 2) return "mesh is not 2:1 balanced at block " + std::to_string(b);
        if (kind == 1) {
          const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
          for (int f2 = 0; f2 < 2; ++f2)
            for (int f1 = 0; f1 < 2; ++f1) {
              int c[3];
              c[ax] = s < 0 ? 1 : 0;
              c[o1] = f1;
              c[o2] = f2;
              push(ox, 1, child(t, c[0], c[1], c[2]), f1, f2);
            }
        } else {
          push(ox, kind, same_or_coarser(t, kind), 0, 0);
        }
      }
    // edges (a, b) = (x1, x2), (x1, x3), (x2, x3): m over b outer, n over a inner
    const int ea[3] = {0, 0, 1}, eb[3] = {1, 2, 2};
    for (int e = 0; e < 3; ++e) {
      const int a = ea[e], bb = eb[e], fr = 3 - a - bb;
      for (int m = -1; m <= 1; m += 2)
        for (int s = -1; s <= 1; s += 2) {
          int ox[3] = {0, 0, 0};
          ox[a] = s;
          ox[bb] = m;
          long long t[3];
          if (!target(ox, t)) continue;
          const int kind = cover(L, t);
          if (kind == 2) return "mesh is not 2:1 balanced at block " + std::to_string(b);
          if (kind == 1) {
            for (int f1 = 0; f1 < 2; ++f1) {
              int c[3];
              c[a] = s < 0 ? 1 : 0;
              c[bb] = m < 0 ? 1 : 0;
              c[fr] = f1;
              push(ox, 1, child(t, c[0], c[1], c[2]), f1, 0);
            }
          } else if (kind == 0 || (my[a] == s && my[bb] == m)) {
            push(ox, kind, same_or_coarser(t, kind), 0, 0);
          }
        }
    }
    // corners: x3 outer, x2, x1 inner
    for (int l = -1; l <= 1; l += 2)
      for (int m = -1; m <= 1; m += 2)
        for (int s = -1; s <= 1; s += 2) {
          int ox[3] = {s, m, l};
          long long t[3];
          if (!target(ox, t)) continue;
          const int kind = cover(L, t);
          if (kind == 2) return "mesh is not 2:1 balanced at block " + std::to_string(b);
          if (kind == 1) {
            push(ox, 1, child(t, s < 0 ? 1 : 0, m < 0 ? 1 : 0, l < 0 ? 1 : 0), 0, 0);
          } else if (kind == 0 || (my[0] == s && my[1] == m && my[2] == l)) {
            push(ox, kind, same_or_coarser(t, kind), 0, 0);
          }
        }
    return "";
  }

  // entry of block q in the list of block r at offset ox (the sender's view)
  const Nbr* entry(int r, int q, const int ox[3]) const {
    for (const Nbr& e : nbrs[r])
      if (e.id == q && e.ox[0] == ox[0] && e.ox[1] == ox[1] && e.ox[2] == ox[2]) return &e;
    return nullptr;
  }
};

// --- the plan ----------------------------------------------------------------
// A vertex reference: blkarr = 2 * block + (0 fine array, 1 coarse buffer),
// p = flat index in that array.
struct Ref {
  int blkarr;
  int p;
};

struct Plan {
  // destinations, interior vertices first (n_interior of them: they are
  // also sources, so their new values go through a scratch), then ghosts
  std::vector<Ref> dst;
  std::vector<long long> off{0};
  std::vector<Ref> src;
  std::vector<double> inv;
  long long n_interior = 0;
  std::vector<std::array<int, 7>> boxes;  // block, si, ei, sj, ej, sk, ek (coarse indices)
  std::vector<int> prolongates;           // per block: has a coarser neighbour
};

// Builds the plan for all blocks of `t` (fine arrays of nx + 9 vertices per
// axis, coarse buffers of nx/2 + 9).  Returns "" or an error message.
inline std::string build_plan(const Tree& t, Plan& P) {
  Axis fa[3], ca[3];
  int fn[3], cn[3], half_f[3], half_l[3];
  for (int d = 0; d < 3; ++d) {
    fa[d] = axis_of(t.nx[d], MNG);
    ca[d] = axis_of(t.nx[d] / 2, MCNG);
    fn[d] = t.nx[d] + 2 * MNG + 1;
    cn[d] = t.nx[d] / 2 + 2 * MCNG + 1;
    half_f[d] = t.nx[d] / 2;
    half_l[d] = t.nx[d] / 2 - MCNG;
  }
  const int fpts = fn[0] * fn[1] * fn[2], cpts = cn[0] * cn[1] * cn[2];
  P.prolongates.assign(t.n, 0);
  Plan G;  // the ghost destinations, appended after the interior ones
  // per destination of the current block: ordered sources, and whether the
  // list starts from the vertex's own value
  std::vector<std::vector<Ref>> lst[2] = {std::vector<std::vector<Ref>>(fpts),
                                          std::vector<std::vector<Ref>>(cpts)};
  std::vector<char> own[2] = {std::vector<char>(fpts), std::vector<char>(cpts)};
  for (int q = 0; q < t.n; ++q) {
    for (int A = 0; A < 2; ++A) {
      const Axis* ax = A ? ca : fa;
      const int* nn = A ? cn : fn;
      for (int k = 0; k < nn[2]; ++k)
        for (int j = 0; j < nn[1]; ++j)
          for (int i = 0; i < nn[0]; ++i) {
            const int p = i + nn[0] * (j + nn[1] * k);
            lst[A][p].clear();
            own[A][p] = i >= ax[0].vs && i <= ax[0].ve && j >= ax[1].vs && j <= ax[1].ve &&
                        k >= ax[2].vs && k <= ax[2].ve;
          }
    }
    int mult[7][7][7];
    for (int k = 0; k < 7; ++k)
      for (int j = 0; j < 7; ++j)
        for (int i = 0; i < 7; ++i) mult[k][j][i] = (i >= 1 && i <= 5 && j >= 1 && j <= 5 && k >= 1 && k <= 5);
    // contributions of one box pair: dst array A of q <- src array SA of block r
    auto add = [&](int A, const Range* dr, int r, int SA, const Range* sr, bool overwrite) {
      const int* dn = A ? cn : fn;
      const int* sn = SA ? cn : fn;
      for (int d = 0; d < 3; ++d)
        if (dr[d].e - dr[d].s != sr[d].e - sr[d].s) return false;
      for (int k = dr[2].s; k <= dr[2].e; ++k)
        for (int j = dr[1].s; j <= dr[1].e; ++j)
          for (int i = dr[0].s; i <= dr[0].e; ++i) {
            const int p = i + dn[0] * (j + dn[1] * k);
            const int si = sr[0].s + (i - dr[0].s), sj = sr[1].s + (j - dr[1].s),
                      sk = sr[2].s + (k - dr[2].s);
            const Ref s{2 * r + SA, si + sn[0] * (sj + sn[1] * sk)};
            if (overwrite) {
              lst[A][p].assign(1, s);
              own[A][p] = 0;
            } else {
              lst[A][p].push_back(s);
            }
          }
      return true;
    };
    for (const Nbr& nb : t.nbrs[q]) {
      const int r = nb.id;
      const int rev[3] = {-nb.ox[0], -nb.ox[1], -nb.ox[2]};
      Range dr[3], sr[3], cr[3];
      bool ok = true;
      if (nb.rel == 0) {
        for (int d = 0; d < 3; ++d) {
          dr[d] = set_same(nb.ox[d], fa[d]);
          sr[d] = load_same(rev[d], fa[d]);
          cr[d] = compact(0, nb.ox[d], 0, 0);
        }
        ok = add(0, dr, r, 0, sr, false);
        if (ok && !t.all_same[q]) {
          for (int d = 0; d < 3; ++d) {
            dr[d] = set_same(nb.ox[d], ca[d]);
            sr[d] = load_same(rev[d], ca[d]);
          }
          ok = add(1, dr, r, 1, sr, false);
        }
      } else if (nb.rel < 0) {
        const Nbr* se = t.entry(r, q, rev);
        if (!se) return "neighbour lists disagree at block " + std::to_string(q);
        for (int d = 0; d < 3; ++d) {
          dr[d] = set_from_coarser(nb.ox[d], ca[d], t.loc[q][d]);
          sr[d] = load_to_finer(rev[d], fa[d], fi_of(*se, d), half_l[d]);
          cr[d] = compact(-1, nb.ox[d], t.loc[q][d], 0);
        }
        ok = add(1, dr, r, 0, sr, true);
        P.prolongates[q] = 1;
      } else {
        for (int d = 0; d < 3; ++d) {
          const int fi = fi_of(nb, d);
          dr[d] = set_from_finer(nb.ox[d], fa[d], fi, half_f[d]);
          sr[d] = load_to_coarser(rev[d], ca[d]);
          cr[d] = compact(1, nb.ox[d], 0, fi);
        }
        ok = add(0, dr, r, 1, sr, false);
      }
      if (!ok) return "message sizes disagree at block " + std::to_string(q);
      for (int k = cr[2].s; k <= cr[2].e; ++k)
        for (int j = cr[1].s; j <= cr[1].e; ++j)
          for (int i = cr[0].s; i <= cr[0].e; ++i) ++mult[k][j][i];
    }
    // destinations: every ghost, and interior vertices in a zone of
    // multiplicity > 1; a coarse buffer only where its block prolongates
    // (coarse ghosts are read by nothing else, and the coarse interior is
    // restricted afresh by every update)
    for (int A = 0; A < 2; ++A) {
      if (A == 1 && !P.prolongates[q]) continue;
      const Axis* ax = A ? ca : fa;
      const int* nn = A ? cn : fn;
      for (int k = 0; k < nn[2]; ++k)
        for (int j = 0; j < nn[1]; ++j)
          for (int i = 0; i < nn[0]; ++i) {
            const int p = i + nn[0] * (j + nn[1] * k);
            const int m = mult[zone_of(k, ax[2])][zone_of(j, ax[1])][zone_of(i, ax[0])];
            if (own[A][p] && m <= 1 && lst[A][p].empty()) continue;
            const bool interior = i >= ax[0].vs && i <= ax[0].ve && j >= ax[1].vs &&
                                  j <= ax[1].ve && k >= ax[2].vs && k <= ax[2].ve;
            Plan& T = interior ? P : G;
            T.dst.push_back(Ref{2 * q + A, p});
            if (own[A][p]) T.src.push_back(Ref{2 * q + A, p});
            for (const Ref& s : lst[A][p]) T.src.push_back(s);
            T.off.push_back((long long)T.src.size());
            T.inv.push_back(m > 1 ? 1.0 / (double)m : 1.0);
          }
    }
    // prolongation boxes of the coarser neighbours, in neighbour order
    for (const Nbr& nb : t.nbrs[q]) {
      if (nb.rel >= 0) continue;
      std::array<int, 7> box;
      box[0] = q;
      for (int d = 0; d < 3; ++d) {
        const int pcng = MNG / 2;
        int s, e;
        if (nb.ox[d] > 0) {
          s = ca[d].ve + 1;
          e = ca[d].ve + pcng;
        } else if (nb.ox[d] < 0) {
          s = ca[d].vs - pcng;
          e = ca[d].vs - 1;
        } else {
          s = ca[d].vs;
          e = ca[d].ve;
          if ((t.loc[q][d] & 1) == 0)
            e += pcng;
          else
            s -= pcng;
        }
        box[1 + 2 * d] = s;
        box[2 + 2 * d] = e;
      }
      P.boxes.push_back(box);
    }
  }
  P.n_interior = (long long)P.dst.size();
  const long long base = (long long)P.src.size();
  P.dst.insert(P.dst.end(), G.dst.begin(), G.dst.end());
  P.src.insert(P.src.end(), G.src.begin(), G.src.end());
  P.inv.insert(P.inv.end(), G.inv.begin(), G.inv.end());
  for (size_t d = 1; d < G.off.size(); ++d) P.off.push_back(base + G.off[d]);
  return "";
}

}  // namespace z4c_mesh
