// FP64 kernels of the MCV step on gfx950.
//
// One RK stage is two launches:
//   stage_kernel     per (panel, tile): stage the tile and its one-cell halo
//                    of (phi, u, v) in LDS with the derived E, phiG, contra
//                    winds; Riemann derivatives at cell ends; interior
//                    derivatives; tendencies; RK update.  Writes the raw
//                    updated interior and the band (a raw copy of the H+1
//                    lines nearest each panel side).
//   boundary_kernel  one item per ghost point, cube-edge point or cube
//                    vertex: ghost interpolation into the halo and edge/vertex
//                    averaging (the reference's correct_bdy_ghost), reading
//                    the band so that no thread reads a point another thread
//                    rewrites.
// phiG is never stored: every state keeps phiG = phi*sqrtG exactly
// (time_scheme_mod.F90:85), so it is recomputed from phi where needed.
#pragma once

#include "mcv_device.hpp"
#include "mcv_stage_fp64.hpp"

namespace mcv {

// Work item t of correct_bdy_ghost on one level, t in [0, boundary_items()):
// one ghost point, one edge point or one cube vertex per item.  Every value is
// read from the band (the raw lines the stage kernel wrote) and the
// line-major matrix tables, contiguously along the lines; the items are
// independent, each a short chain of loads.
template <int DOF>
__device__ __forceinline__ void boundary_item(const MeshDev &m, const BoundaryArgs &a,
                                              const int t) {
  constexpr int H = DOF - 1;
  const int npv = m.npv, nh = m.nh;
  const double dx = m.dx;
  const int nghost = 6 * 4 * H * npv, nedge = 12 * npv;
  if (t < nghost) {
    // (panel, side, depth k, along-edge s), s fastest
    const int s = t % npv + 1, r = t / npv, k = r % H + 1, ps = r / H;
    const int side = ps % 4, p = ps / 4 + 1;
    const HaloSrc hs = kHalo[p - 1][side];
    // Ghost point at depth k (ghost_mod.f90:154-227): the neighbour's line at
    // normal index k+1 (or nPV-k) is band depth k+1 of the side it is parallel
    // to; sphere winds through P2SP there and back through SP2P at the ghost
    // point (spatial_operators_mod.F90:408-492).
    const int src = hs.col ? (hs.far ? 1 : 0) : (hs.far ? 3 : 2);
    const int c = m.giref[(k - 1) * npv + (s - 1)];
    const double cf = m.gcoef[(k - 1) * npv + (s - 1)];
    double qp[DOF], qz[DOF], qm[DOF];
#pragma unroll
    for (int l = 0; l < DOF; ++l) {
      const int tpos = (c - 1) * H + 1 + l;
      const int al = hs.rev ? npv + 1 - tpos : tpos;
      const double ph = ldg(a.band_raw, band_slot(npv, H, hs.q, src, k + 1, 0, al));
      const double u = ldg(a.band_raw, band_slot(npv, H, hs.q, src, k + 1, 1, al));
      const double v = ldg(a.band_raw, band_slot(npv, H, hs.q, src, k + 1, 2, al));
      qp[l] = ph;
      apply2(m.p2sp_band + 4 * band_point(npv, H, hs.q, src, k + 1, al), u, v, qz[l], qm[l]);
    }
    int hi, hj;
    if (side == 0) { hi = 1 - k; hj = s; }
    else if (side == 1) { hi = npv + k; hj = s; }
    else if (side == 2) { hi = s; hj = 1 - k; }
    else { hi = s; hj = npv + k; }
    const long long g = pidx(nh, H, p, hi, hj);
    const double zh = ghost_interp<DOF>(qz, cf, dx), mh = ghost_interp<DOF>(qm, cf, dx);
    const double ph = ghost_interp<DOF>(qp, cf, dx);
    double uh, vh;
    apply2(m.sp2p_ghost + 4 * ghost_point(npv, H, p, side, k, s), zh, mh, uh, vh);
    a.pu[g] = D2{ph, uh};
    a.v[g] = vh;
  } else if (t < nghost + nedge) {
    // Edge averaging (spatial_operators_mod.F90:298-345), vertices excluded.
    const int s = (t - nghost) % npv + 1, e = (t - nghost) / npv;
    if (s > 1 && s < npv) {
      const EdgePair ep = kEdges[e];
      const int sb = ep.rev ? npv + 1 - s : s;
      const double pa = ldg(a.band_raw, band_slot(npv, H, ep.pa, ep.sa, 1, 0, s));
      const double ua = ldg(a.band_raw, band_slot(npv, H, ep.pa, ep.sa, 1, 1, s));
      const double va = ldg(a.band_raw, band_slot(npv, H, ep.pa, ep.sa, 1, 2, s));
      const double pb = ldg(a.band_raw, band_slot(npv, H, ep.pb, ep.sb, 1, 0, sb));
      const double ub = ldg(a.band_raw, band_slot(npv, H, ep.pb, ep.sb, 1, 1, sb));
      const double vb = ldg(a.band_raw, band_slot(npv, H, ep.pb, ep.sb, 1, 2, sb));
      const uint32_t ba = band_point(npv, H, ep.pa, ep.sa, 1, s);
      const uint32_t bb = band_point(npv, H, ep.pb, ep.sb, 1, sb);
      int ia, ja, ib, jb;
      edge_ij(npv, ep.sa, s, ia, ja);
      edge_ij(npv, ep.sb, sb, ib, jb);
      const long long ga = pidx(nh, H, ep.pa, ia, ja), gb = pidx(nh, H, ep.pb, ib, jb);
      double za, ma, zb, mb;
      apply2(m.p2sp_band + 4 * ba, ua, va, za, ma);
      apply2(m.p2sp_band + 4 * bb, ub, vb, zb, mb);
      const double ph = 0.5 * (pa + pb), zu = 0.5 * (za + zb), mu = 0.5 * (ma + mb);
      double u1, v1, u2, v2;
      apply2(m.sp2p_edge + 4 * edge_point(npv, ep.pa, ep.sa, s), zu, mu, u1, v1);
      apply2(m.sp2p_edge + 4 * edge_point(npv, ep.pb, ep.sb, sb), zu, mu, u2, v2);
      a.pu[ga] = D2{ph, u1}; a.v[ga] = v1;
      a.pu[gb] = D2{ph, u2}; a.v[gb] = v2;
    }
  } else {
    // Cube vertices (spatial_operators_mod.F90:347-386), reference summation order.
    const int vx = t - nghost - nedge;
    double ph[3], z[3], mr[3];
    long long gg[3];
    uint32_t eg[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const VtxPt vp = kVerts[vx][l];
      const int i = vp.ci ? npv : 1, j = vp.cj ? npv : 1;
      const int side = vp.ci ? 1 : 0;  // either edge through the corner holds its raw value
      ph[l] = ldg(a.band_raw, band_slot(npv, H, vp.p, side, 1, 0, j));
      const double u = ldg(a.band_raw, band_slot(npv, H, vp.p, side, 1, 1, j));
      const double v = ldg(a.band_raw, band_slot(npv, H, vp.p, side, 1, 2, j));
      gg[l] = pidx(nh, H, vp.p, i, j);
      eg[l] = edge_point(npv, vp.p, side, j);
      apply2(m.p2sp_band + 4 * band_point(npv, H, vp.p, side, 1, j), u, v, z[l], mr[l]);
    }
    const double pv = (ph[0] + ph[1] + ph[2]) / 3.0;
    const double zv = (z[0] + z[1] + z[2]) / 3.0;
    const double mv = (mr[0] + mr[1] + mr[2]) / 3.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double u, v;
      apply2(m.sp2p_edge + 4 * eg[l], zv, mv, u, v);
      a.pu[gg[l]] = D2{pv, u};
      a.v[gg[l]] = v;
    }
  }
}

template <int DOF>
__global__ __launch_bounds__(256) void boundary_kernel(MeshDev m, BoundaryArgs a) {
  const int nitems = boundary_items(DOF, m.npv);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nitems; t += gridDim.x * blockDim.x)
    boundary_item<DOF>(m, a, t);
}

// Reference-layout copy-out of one level: phi, phiG = phi*sqrtG, u, v and
// the contravariant winds (or the host's initial ones before the first step).
static __global__ void pack_kernel(MeshDev m, const double *phi, const double *u, const double *v,
                            const double *cu0, const double *cv0, double *o_phi,
                            double *o_phiG, double *o_u, double *o_v, double *o_cu,
                            double *o_cv) {
  const long long np = 6LL * m.nh * m.nh;
  const long long per = (long long)m.nh * m.nh;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    const long long p = g / per, off = g - p * per;
    const long long mi = p * m.metric_pstride + off;
    const double ph = phi[g], uu = u[g], vv = v[g];
    o_phi[g] = ph;
    o_phiG[g] = ph * m.sqrtG[mi];
    o_u[g] = uu;
    o_v[g] = vv;
    if (cu0) {
      o_cu[g] = cu0[g];
      o_cv[g] = cv0[g];
    } else {
      o_cu[g] = m.ig11[mi] * uu + m.ig12[mi] * vv;
      o_cv[g] = m.ig12[mi] * uu + m.ig22[mi] * vv;
    }
  }
}

// Sphere winds on every point (spatial_operators_mod.F90:391-405).
static __global__ void sphere_winds_kernel(MeshDev m, const double *u, const double *v, double *zon,
                                    double *mer) {
  const long long np = 6LL * m.nh * m.nh;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    double z, w;
    apply2(m.p2sp + 4 * g, u[g], v[g], z, w);
    zon[g] = z;
    mer[g] = w;
  }
}

// {a, b} pairs of a time level into plain arrays (read-back paths; b may be
// null).
static __global__ void unpack_pairs_kernel(long long n, const D2 *pu, double *a, double *b) {
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x) {
    const D2 x = pu[g];
    a[g] = x.a;
    if (b) b[g] = x.b;
  }
}

// Plain arrays a, b into {a, b} pairs (the layout of a time level).
static __global__ void pack_pairs_kernel(long long n, const double *a, const double *b, D2 *pu) {
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x)
    pu[g] = D2{a[g], b[g]};
}

// FillValue of constants_mod.F90:20.
constexpr double kFillValue = -9999999999999999.0;

// Point fields of one history record, as the output block of MCV_SW.F90:62-70
// leaves them in stat(new) after a time step:
//   addFillValue (output_mod.F90:230-248) sets u, v on the four halo corners
//   to FillValue; fill_ghost (ghost_mod.f90:160-179) left phi there at
//   FillValue; convert_wind_P2SP (spatial_operators_mod.F90:391-405) then
//   forms the sphere winds on every point, corners included; contraU/V are
//   cov2contrav (projection_mod.F90:138-148) of the same u, v, as the last
//   boundary pass formed them; phi_t = phi + phi_s (output_mod.F90:207).
static __global__ void history_point_kernel(MeshDev m, const double *phi, const double *u,
                                            const double *v, double *o_u, double *o_v,
                                            double *o_cu, double *o_cv, double *o_phi,
                                            double *o_phit, double *o_zon, double *o_mer) {
  const int H = m.dof - 1, npv = m.npv, nh = m.nh;
  const long long per = (long long)nh * nh, np = 6 * per;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    const long long p = g / per, off = g - p * per;
    const int i = (int)(off % nh) + 1 - H, j = (int)(off / nh) + 1 - H;
    const bool corner = (i < 1 || i > npv) && (j < 1 || j > npv);
    const long long mi = p * m.metric_pstride + off;
    const double uu = corner ? kFillValue : u[g];
    const double vv = corner ? kFillValue : v[g];
    const double ph = corner ? kFillValue : phi[g];
    o_u[g] = uu;
    o_v[g] = vv;
    o_cu[g] = m.ig11[mi] * uu + m.ig12[mi] * vv;
    o_cv[g] = m.ig12[mi] * uu + m.ig22[mi] * vv;
    o_phi[g] = ph;
    o_phit[g] = ph + m.phi_s[p * m.phis_pstride + off];
    double z, w;
    apply2(m.p2sp + 4 * g, uu, vv, z, w);
    o_zon[g] = z;
    o_mer[g] = w;
  }
}

// calc_VIA (diag_mod.F90:69-93) of three point fields: cell (i, j) of panel p
// is the weighted sum of its DOF x DOF points (pvIdx, mesh_mod.F90:63-73),
// accumulated from 0 with iDOF fastest, the weights those of
// mesh_mod.F90:183-203 ([1,4,1] x [1,4,1] / 36, [1,3,3,1] x [1,3,3,1] / 80).
// Cells are (n + 2)^2 per panel, the halo ring (ics = 0, ice = n + 1) zero.
template <int DOF>
__global__ void cell_average_kernel(MeshDev m, int n, const double *a, const double *b,
                                    const double *c, double *oa, double *ob, double *oc) {
  constexpr int H = DOF - 1;
  constexpr double kRow[4] = {1.0, DOF == 3 ? 4.0 : 3.0, DOF == 3 ? 1.0 : 3.0, 1.0};
  constexpr double kDen = DOF == 3 ? 36.0 : 80.0;
  const int nc = n + 2, nh = m.nh;
  const long long per = (long long)nc * nc, ncell = 6 * per;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ncell;
       t += (long long)gridDim.x * blockDim.x) {
    const long long p = t / per, off = t - p * per;
    const int ic = (int)(off % nc), jc = (int)(off / nc);  // 0 .. n+1
    double sa = 0.0, sb = 0.0, sc = 0.0;
    if (ic >= 1 && ic <= n && jc >= 1 && jc <= n) {
#pragma unroll
      for (int jd = 0; jd < DOF; ++jd)
#pragma unroll
        for (int id = 0; id < DOF; ++id) {
          const double w = (kRow[id] * kRow[jd]) / kDen;
          const int ip = (ic - 1) * H + id + 1, jp = (jc - 1) * H + jd + 1;
          const long long g = pidx(nh, H, (int)p + 1, ip, jp);
          sa = sa + w * a[g];
          sb = sb + w * b[g];
          sc = sc + w * c[g];
        }
    }
    oa[t] = sa;
    ob[t] = sb;
    oc[t] = sc;
  }
}

// ---------------------------------------------------------------------------
// FP64 integrals (diag_mod.F90:10-67): per-cell sums in the reference order,
// then a deterministic double-double tree over cells and blocks.
// ---------------------------------------------------------------------------
struct DD { double hi, lo; };

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  // two-sum of the high parts, then fold the low parts in (Knuth)
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (b.hi - bb);
  const double lo = err + a.lo + b.lo;
  const double hi = s + lo;
  return DD{hi, lo - (hi - s)};
}

constexpr int kDiagThreads = 256;

template <int DOF>
__global__ __launch_bounds__(kDiagThreads) void diag_partial_kernel(MeshDev m, const double *phi,
                                                                    const double *u,
                                                                    const double *v,
                                                                    const double *cu0,
                                                                    const double *cv0,
                                                                    double *partials) {
  constexpr int H = DOF - 1;
  const int N = m.n, nh = m.nh;
  const long long ncell = 6LL * N * N;
  double w1[DOF];
  if constexpr (DOF == 3) { w1[0] = 1.0; w1[1] = 4.0; w1[2] = 1.0; }
  else { w1[0] = 1.0; w1[1] = 3.0; w1[2] = 3.0; w1[3] = 1.0; }
  const double norm = DOF == 3 ? 36.0 : 80.0;  // mesh_mod.F90:183-203
  DD ms{0.0, 0.0}, es{0.0, 0.0};
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(c / ((long long)N * N)) + 1;
    const int rem = (int)(c - (long long)(p - 1) * N * N);
    const int cj = rem / N + 1, ci = rem - (cj - 1) * N + 1;
    double mc = 0.0, ec = 0.0;
    for (int b = 0; b < DOF; ++b)
      for (int aa = 0; aa < DOF; ++aa) {
        const int i = (ci - 1) * H + 1 + aa, j = (cj - 1) * H + 1 + b;
        const long long g = pidx(nh, H, p, i, j);
        const long long off = g - (long long)(p - 1) * nh * nh;
        const long long mi = (long long)(p - 1) * m.metric_pstride + off;
        const double w = (w1[aa] * w1[b]) / norm;
        const double sg = m.sqrtG[mi];
        const double ph = phi[g], uu = u[g], vv = v[g];
        const double cu = cu0 ? cu0[g] : m.ig11[mi] * uu + m.ig12[mi] * vv;
        const double cv = cv0 ? cv0[g] : m.ig12[mi] * uu + m.ig22[mi] * vv;
        const double ps = m.phi_s[(long long)(p - 1) * m.phis_pstride + off];
        mc = mc + w * (ph * sg);
        const double KE = 0.5 * ph * (uu * cu + vv * cv);
        const double PE = 0.5 * ((ph + ps) * (ph + ps));
        ec = ec + w * sg * (KE + PE);
      }
    ms = dd_add(ms, DD{mc, 0.0});
    es = dd_add(es, DD{ec, 0.0});
  }
  __shared__ DD sm[kDiagThreads], se[kDiagThreads];
  sm[threadIdx.x] = ms;
  se[threadIdx.x] = es;
  __syncthreads();
  for (int w = (int)blockDim.x / 2; w > 0; w >>= 1) {  // blockDim.x: a power of two <= kDiagThreads
    if ((int)threadIdx.x < w) {
      sm[threadIdx.x] = dd_add(sm[threadIdx.x], sm[threadIdx.x + w]);
      se[threadIdx.x] = dd_add(se[threadIdx.x], se[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[4 * blockIdx.x + 0] = sm[0].hi;
    partials[4 * blockIdx.x + 1] = sm[0].lo;
    partials[4 * blockIdx.x + 2] = se[0].hi;
    partials[4 * blockIdx.x + 3] = se[0].lo;
  }
}

// The block partials (at most kDiagMaxBlocks) are staged in LDS by the whole
// workgroup, then summed by one thread in block order.
constexpr int kDiagMaxBlocks = 1024;
static __global__ void diag_final_kernel(const double *partials, int nblocks, double *out) {
  __shared__ double sp[4 * kDiagMaxBlocks];
  for (int k = threadIdx.x; k < 4 * nblocks; k += blockDim.x) sp[k] = partials[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    DD ms{0.0, 0.0}, es{0.0, 0.0};
    for (int b = 0; b < nblocks; ++b) {
      ms = dd_add(ms, DD{sp[4 * b], sp[4 * b + 1]});
      es = dd_add(es, DD{sp[4 * b + 2], sp[4 * b + 3]});
    }
    out[0] = ms.hi + ms.lo;
    out[1] = es.hi + es.lo;
  }
}

}  // namespace mcv
