// Host-side mesh preparation: validation, reverse Cuthill-McKee renumbering,
// structure-of-arrays packing and upload.
//
// Follows the reference mesh conventions (jax_fvm/src/mesh.py:28-83,
// 223-321): local edge k = (v_k, v_k+1), neighbours -1 on the boundary,
// face markers 0 interior / 1 periodic / 2 wall / 3 inlet / 4 outlet /
// 5 subsonic inlet. All float arithmetic here is IEEE single precision with
// no contraction (built with -ffp-contract=off), so the precomputed face
// offsets equal what the reference computes inside its jitted residual.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "esr_internal.h"

namespace esr {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& s) { g_last_error = s; }
const char* last_error_cstr() { return g_last_error.c_str(); }

// Reverse Cuthill-McKee over the cell adjacency graph, the algorithm of
// scipy.sparse.csgraph.reverse_cuthill_mckee: components are seeded in
// `seed_order` (a permutation of the cells; NULL = by degree, ties by
// index); breadth-first levels; the unvisited neighbours of each node are
// appended in increasing index, then stably sorted by degree; the final
// order is reversed. perm[new] = old.
std::vector<int32_t> rcm_order(int32_t n, const int32_t* nb, const int32_t* seed_order) {
  std::vector<int32_t> deg(n, 0);
  std::vector<int32_t> adj(3 * (size_t)n, -1);
  for (int32_t i = 0; i < n; ++i) {
    int32_t cnt = 0;
    int32_t tmp[3];
    for (int k = 0; k < 3; ++k) {
      int32_t j = nb[3 * (size_t)i + k];
      if (j < 0 || j == i) continue;
      bool dup = false;
      for (int q = 0; q < cnt; ++q) dup |= (tmp[q] == j);
      if (!dup) tmp[cnt++] = j;
    }
    std::sort(tmp, tmp + cnt);
    for (int q = 0; q < cnt; ++q) adj[3 * (size_t)i + q] = tmp[q];
    deg[i] = cnt;
  }
  std::vector<int32_t> seeds(n);
  if (seed_order) {
    std::vector<char> seen(n, 0);
    for (int32_t q = 0; q < n; ++q) {
      const int32_t s = seed_order[q];
      if (s < 0 || s >= n || seen[s]) throw Error(ESR_ERR_VALUE, "RCM seed order is not a permutation of the cells");
      seen[s] = 1;
      seeds[q] = s;
    }
  } else {
    std::iota(seeds.begin(), seeds.end(), 0);
    std::stable_sort(seeds.begin(), seeds.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
  }
  std::vector<char> vis(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  for (int32_t s : seeds) {
    if (vis[s]) continue;
    vis[s] = 1;
    order.push_back(s);
    size_t head = order.size() - 1;
    while (head < order.size()) {
      int32_t i = order[head++];
      int32_t cand[3];
      int cnt = 0;
      for (int q = 0; q < deg[i]; ++q) {
        int32_t j = adj[3 * (size_t)i + q];
        if (!vis[j]) cand[cnt++] = j;
      }
      std::stable_sort(cand, cand + cnt, [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
      for (int q = 0; q < cnt; ++q) {
        vis[cand[q]] = 1;
        order.push_back(cand[q]);
      }
    }
    if ((int32_t)order.size() == n) break;
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// blend^3 of helper.py:262-295 per (renumbered) cell, float32.
std::vector<float> sponge_weights(const esr_mesh* m, float width) {
  std::vector<float> g(m->n);
  for (int32_t i = 0; i < m->n; ++i) {
    float y = m->bary_y[i];
    float d1 = y - m->y_min, d2 = m->y_max - y;
    float dist = d1 < d2 ? d1 : d2;
    float b = (width - dist) / width;
    b = b < 0.f ? 0.f : (b > 1.f ? 1.f : b);
    float b2 = b * b;
    g[i] = b2 * b;
  }
  return g;
}

esr_mesh* build_mesh(const esr_mesh_desc* d) {
  if (!d) throw Error(ESR_ERR_VALUE, "null mesh description");
  const int32_t n = d->n_cells, nf = d->n_faces, np = d->n_points;
  if (n <= 0 || nf <= 0 || np <= 0) throw Error(ESR_ERR_VALUE, "empty mesh");
  if ((int64_t)n > (int64_t)FREF_BIDX) throw Error(ESR_ERR_VALUE, "too many cells for one mesh");
  if (!d->points || !d->tris || !d->neighbors || !d->faces || !d->face_markers ||
      !d->face_connectivity || !d->face_connectivity_opposite || !d->barycenter || !d->area ||
      !d->surface || !d->normals)
    throw Error(ESR_ERR_VALUE, "mesh description has a null array");

  auto* m = new esr_mesh();
  try {
    m->n = n;
    m->n_faces = nf;
    m->n_points = np;
    m->sponge = d->sponge ? 1 : 0;
    const int32_t* NB = d->neighbors;
    const int32_t* FC = d->face_connectivity;
    const int32_t* FCO = d->face_connectivity_opposite;
    const int32_t* FM = d->face_markers;

    // validation
    for (int64_t t = 0; t < 3 * (int64_t)n; ++t) {
      if (FC[t] < 0 || FC[t] >= nf) throw Error(ESR_ERR_VALUE, "face_connectivity out of range");
      if (NB[t] < -1 || NB[t] >= n) throw Error(ESR_ERR_VALUE, "neighbors out of range");
    }
    for (int32_t f = 0; f < nf; ++f)
      if (FCO[f] < 0 || FCO[f] >= nf) throw Error(ESR_ERR_VALUE, "face_connectivity_opposite out of range");

    // renumbering
    if (d->renumber) {
      m->perm = rcm_order(n, NB, d->rcm_seed_order);
    } else {
      m->perm.resize(n);
      std::iota(m->perm.begin(), m->perm.end(), 0);
    }
    m->iperm.assign(n, -1);
    for (int32_t i = 0; i < n; ++i) m->iperm[m->perm[i]] = i;

    // boundary kind per marker
    std::vector<int32_t> walls;
    for (int32_t q = 0; q < d->n_wall_markers; ++q) walls.push_back(d->wall_markers[q]);
    if (d->n_wall_markers <= 0) walls.push_back(2);
    auto ghost_kind = [&](int32_t mk) -> uint32_t {
      if (mk == 5) return G_SUBSONIC;
      if (mk == 4) return G_OUTLET;
      if (mk == 3) return G_INLET;
      if (std::find(walls.begin(), walls.end(), mk) != walls.end()) return G_WALL;
      return G_WRAP;
    };
    auto slot_of = [&](int32_t cell, int32_t face) -> int32_t {
      for (int k = 0; k < 3; ++k)
        if (FC[3 * (size_t)cell + k] == face) return k;
      return -1;
    };

    const float* P = d->points;
    std::vector<uint32_t> fref(3 * (size_t)n);
    std::vector<float2> bary(n), dxm(3 * (size_t)n), nrm(3 * (size_t)n);
    std::vector<float> len(3 * (size_t)n), area(n);
    m->bary_y.resize(n);
    for (int32_t i = 0; i < n; ++i) {
      const int32_t o = m->perm[i];
      const float bx = d->barycenter[2 * (size_t)o], by = d->barycenter[2 * (size_t)o + 1];
      bary[i] = make_float2(bx, by);
      m->bary_y[i] = by;
      area[i] = d->area[o];
      for (int k = 0; k < 3; ++k) {
        const int32_t f = FC[3 * (size_t)o + k];
        const int32_t mk = FM[f];
        const int32_t j = NB[3 * (size_t)o + k];
        const int32_t a = d->faces[2 * (size_t)f], b = d->faces[2 * (size_t)f + 1];
        if (a < 0 || a >= np || b < 0 || b >= np) throw Error(ESR_ERR_VALUE, "faces out of range");
        // midpoint as jnp.mean of the two vertices, then minus the barycentre
        const float mx = (P[2 * (size_t)a] + P[2 * (size_t)b]) / 2.0f;
        const float my = (P[2 * (size_t)a + 1] + P[2 * (size_t)b + 1]) / 2.0f;
        const size_t s = (size_t)k * n + i;
        dxm[s] = make_float2(mx - bx, my - by);
        nrm[s] = make_float2(d->normals[6 * (size_t)o + 2 * k], d->normals[6 * (size_t)o + 2 * k + 1]);
        len[s] = d->surface[f];
        uint32_t w;
        if (mk >= 2) {
          const uint32_t kind = ghost_kind(mk);
          const int32_t jj = (j >= 0) ? m->iperm[j] : m->iperm[n - 1];
          w = FREF_BND | (kind << FREF_KIND_SHIFT) | (uint32_t)jj;
        } else {
          if (j < 0)
            throw Error(ESR_ERR_VALUE, mk == 1 ? "periodic face without a paired cell"
                                               : "boundary face with marker 0");
          const int32_t fo = (mk == 1) ? FCO[f] : f;
          const int32_t kk = slot_of(j, fo);
          if (kk < 0) throw Error(ESR_ERR_VALUE, "neighbour does not own the shared face");
          w = ((uint32_t)kk << 30) | (mk == 1 ? FREF_PER : 0u) | (uint32_t)m->iperm[j];
        }
        fref[s] = w;
      }
    }

    // least-squares geometry (helper.py:136-153): offsets dl_k (neighbour
    // barycentre, or twice the face-midpoint offset on marker > 0 faces),
    // weights 1/|dl|^2 and the inverse of sum_k w dl dl^T, in double.
    std::vector<float2> wdl(3 * (size_t)n);
    std::vector<float4> ainv(n);
    std::vector<float> iarea(n);
    for (int32_t i = 0; i < n; ++i) {
      double A00 = 0, A01 = 0, A11 = 0;
      for (int k = 0; k < 3; ++k) {
        const size_t s = (size_t)k * n + i;
        float2 dl;
        if (fref_is_bnd(fref[s]) || (fref[s] & FREF_PER)) {
          dl = make_float2(2.0f * dxm[s].x, 2.0f * dxm[s].y);
        } else {
          const float2 bj = bary[fref[s] & FREF_IDX];
          dl = make_float2(bj.x - bary[i].x, bj.y - bary[i].y);
        }
        const double x = dl.x, y = dl.y, w = 1.0 / (x * x + y * y);
        wdl[s] = make_float2((float)(w * x), (float)(w * y));
        A00 += w * x * x;
        A01 += w * x * y;
        A11 += w * y * y;
      }
      const double det = A00 * A11 - A01 * A01;
      if (!(std::fabs(det) > 0.0)) throw Error(ESR_ERR_VALUE, "degenerate least-squares stencil");
      ainv[i] = make_float4((float)(A11 / det), (float)(-A01 / det), (float)(-A01 / det), (float)(A00 / det));
      iarea[i] = (float)(1.0 / (double)area[i]);
    }

    // logical faces. The owner of a shared face is the side whose ORIGINAL
    // cell index is lower (then the lower slot), so the arithmetic of every
    // face is independent of the renumbering; faces are stored in renumbered
    // owner order for locality.
    auto owns = [&](int32_t i, int k) -> bool {
      const uint32_t w = fref[(size_t)k * n + i];
      if (fref_is_bnd(w)) return true;
      const int32_t j = (int32_t)(w & FREF_IDX);
      const int kk = (int)(w >> 30);
      const int32_t oi = m->perm[i], oj = m->perm[j];
      return oi < oj || (oi == oj && k < kk);
    };
    std::vector<uint32_t> cface(3 * (size_t)n, 0xffffffffu), f_own, f_ref;
    std::vector<float2> f_nrm;
    std::vector<float> f_len;
    for (int32_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        if (!owns(i, k)) continue;
        const size_t s = (size_t)k * n + i;
        const uint32_t id = (uint32_t)f_own.size();
        f_own.push_back(((uint32_t)k << 30) | (uint32_t)i);
        f_ref.push_back(fref[s]);
        f_nrm.push_back(nrm[s]);
        f_len.push_back(len[s]);
        if (cface[s] != 0xffffffffu) throw Error(ESR_ERR_VALUE, "inconsistent face pairing");
        cface[s] = id;
        if (!fref_is_bnd(fref[s])) {
          const uint32_t j = fref[s] & FREF_IDX, kk = fref[s] >> 30;
          const size_t sp = (size_t)kk * n + j;
          if (cface[sp] != 0xffffffffu || owns((int32_t)j, (int)kk))
            throw Error(ESR_ERR_VALUE, "inconsistent face pairing");
          cface[sp] = id | CFACE_NEG;
        }
      }
    }
    for (size_t s = 0; s < cface.size(); ++s)
      if (cface[s] == 0xffffffffu) throw Error(ESR_ERR_VALUE, "a cell side has no logical face");
    const int32_t nlf = (int32_t)f_own.size();

    // sponge extremes from all mesh points (helper.py:271-272)
    float ymin = P[1], ymax = P[1];
    for (int32_t p = 1; p < np; ++p) {
      float y = P[2 * (size_t)p + 1];
      ymin = y < ymin ? y : ymin;
      ymax = y > ymax ? y : ymax;
    }
    m->y_min = ymin;
    m->y_max = ymax;
    const float Ly = ymax - ymin;
    std::vector<float> g = sponge_weights(m, 0.2f * Ly);

    // force-marker faces: first owner cell, first matching slot (helper.py:370-380)
    const int32_t fmk = d->force_marker;
    std::vector<int32_t> first_owner(nf, -1);
    for (int32_t c = 0; c < n; ++c)
      for (int k = 0; k < 3; ++k) {
        int32_t f = FC[3 * (size_t)c + k];
        if (first_owner[f] < 0) first_owner[f] = c;
      }
    std::vector<int32_t> wcell;
    std::vector<float> wnx, wny, wds;
    for (int32_t f = 0; f < nf; ++f) {
      if (FM[f] != fmk) continue;
      int32_t c = first_owner[f];
      if (c < 0) throw Error(ESR_ERR_VALUE, "force face owned by no cell");
      int32_t k = slot_of(c, f);
      wcell.push_back(m->iperm[c]);
      wnx.push_back(d->normals[6 * (size_t)c + 2 * k]);
      wny.push_back(d->normals[6 * (size_t)c + 2 * k + 1]);
      wds.push_back(d->surface[f]);
    }

    upload(m->b_fref, fref);
    upload(m->b_dxm, dxm);
    upload(m->b_wdl, wdl);
    upload(m->b_ainv, ainv);
    upload(m->b_nrm, nrm);
    upload(m->b_fown, f_own);
    upload(m->b_fref2, f_ref);
    upload(m->b_fnrm, f_nrm);
    upload(m->b_flen, f_len);
    upload(m->b_cface, cface);
    upload(m->b_iarea, iarea);
    upload(m->b_sponge, g);
    upload(m->b_len, len);
    upload(m->b_area, area);
    upload(m->b_perm, m->perm);
    upload(m->b_wcell, wcell);
    upload(m->b_wnx, wnx);
    upload(m->b_wny, wny);
    upload(m->b_wds, wds);

    MeshDev& md = m->dev;
    md.n = n;
    md.nf = nlf;
    md.n_wall = (int32_t)wcell.size();
    md.fref = m->b_fref.as<uint32_t>();
    md.dxm = m->b_dxm.as<float2>();
    md.wdl = m->b_wdl.as<float2>();
    md.ainv = m->b_ainv.as<float4>();
    md.nrm = m->b_nrm.as<float2>();
    md.f_own = m->b_fown.as<uint32_t>();
    md.f_ref = m->b_fref2.as<uint32_t>();
    md.f_nrm = m->b_fnrm.as<float2>();
    md.f_len = m->b_flen.as<float>();
    md.cface = m->b_cface.as<uint32_t>();
    md.inv_area = m->b_iarea.as<float>();
    md.sponge_g = m->b_sponge.as<float>();
    md.len = m->b_len.as<float>();
    md.area = m->b_area.as<float>();
    md.wall_cell = m->b_wcell.as<int32_t>();
    md.wall_nx = m->b_wnx.as<float>();
    md.wall_ny = m->b_wny.as<float>();
    md.wall_ds = m->b_wds.as<float>();
    const float* is = d->inlet_subsonic;
    md.inlet_subsonic = is ? make_float4(is[0], is[1], is[2], is[3]) : make_float4(1.f, 1.f, 0.f, 1.f);
    m->device_bytes = (int64_t)(m->b_fref.bytes + m->b_dxm.bytes + m->b_wdl.bytes + m->b_ainv.bytes +
                                m->b_nrm.bytes + m->b_fown.bytes + m->b_fref2.bytes + m->b_fnrm.bytes +
                                m->b_flen.bytes + m->b_cface.bytes + m->b_iarea.bytes + m->b_sponge.bytes +
                                m->b_perm.bytes + m->b_wcell.bytes + m->b_wnx.bytes + m->b_wny.bytes +
                                m->b_wds.bytes + m->b_len.bytes + m->b_area.bytes);
  } catch (...) {
    delete m;
    throw;
  }
  return m;
}

}  // namespace esr

esr_mesh::~esr_mesh() {
  if (ws) esr::free_workspace(ws);
}
