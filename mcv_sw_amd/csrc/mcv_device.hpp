// Device-side building blocks of the MCV step: MCV reconstruction formulas,
// the Riemann flux, the cubed-sphere connectivity tables and the work
// descriptors passed to the kernels.
//
// FP64 arithmetic here keeps the reference's operand order, and the library
// is compiled with -ffp-contract=off, so the FP64 path rounds exactly where
// the reference Fortran rounds (division and sqrt are IEEE-correct on gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcv {

// ---------------------------------------------------------------------------
// Correctly rounded division by a divisor b whose reciprocal y = RN(1/b) is
// known: q = RN(a*y), r = a - b*q (exact through the FMA), RN(q + r*y) is
// RN(a/b) (Markstein, IBM J. Res. Dev. 34(1), 1990; y within half an ulp of
// 1/b, q within an ulp of a/b).  Valid away from overflow and underflow.
// The result takes the sign of q: for a != 0 the quotient and q are nonzero
// with the sign of a*b, and when r == 0 (q exact) RN(q + r*y) has the
// magnitude of q, so copysign returns q itself, including the sign of a zero
// quotient (a = -0), which the final FMA alone would lose.  Four instructions
// (mul, fma, fma, bit insert) instead of the ~10 of a general IEEE division.
// ---------------------------------------------------------------------------
struct Recip {
  double b, y;  // divisor and RN(1/b)
};

__host__ __device__ inline Recip make_recip(double b) { return Recip{b, 1.0 / b}; }

__device__ __forceinline__ double div_by(double a, Recip d) {
  const double q = a * d.y;
  const double r = fma(-q, d.b, a);
  return copysign(fma(r, d.y, q), q);
}

// Divisors of the MCV formulas, one set per resolution.
struct MCVDivisors {
  Recip dx;   // dh of polyfit (MCV3), fxVIA
  Recip dx2;  // 2*dh of polyfit (MCV4)
  Recip dxx;  // dx**2 of fxxc (MCV4)
};

// ---------------------------------------------------------------------------
// Per-resolution geometry and per-instance state, as the kernels see them.
// ---------------------------------------------------------------------------

// Static mesh data of one resolution on the device.  Arrays are in the
// reference layout (nh*nh points per panel, i fastest, halo included).  The
// metric terms sqrtG and IG are identical on all six panels of the reference
// mesh; metric_pstride is 0 then (one copy serves every panel) and nh*nh
// otherwise.  The same holds for phi_s with phis_pstride.
// Two doubles loaded together (one 16-byte load); two floats (one 8-byte load).
struct alignas(16) D2 {
  double a, b;
};
struct alignas(8) F2 {
  float a, b;
};
// Three floats as one 12-byte record (one dwordx3 load or store; 4-byte
// alignment is all the instruction needs): an FP32C point's {phi', u, v}.
struct alignas(4) F3 {
  float a, b, c;
};
// Four floats loaded together (one 16-byte load).
struct alignas(16) F4 {
  float a, b, c, d;
};

struct MeshDev {
  int dof, n, npv, nh;
  double dx;                 // element width, radians
  MCVDivisors dv;            // dx, 2dx, dx^2 with their reciprocals
  long long metric_pstride;  // panel stride of sqrtG, rsqrtG, ig11, ig12, ig22
  long long phis_pstride;    // panel stride of phi_s
  const double *sqrtG, *rsqrtG, *ig11, *ig12, *ig22;  // rsqrtG = RN(1/sqrtG)
  // The same metric terms packed per point for the stage kernel, so that a
  // point's terms come in one or two 16-byte loads (metric_pstride applies):
  const D2 *mpack;  // 2 per point: {sqrtG, IG11}, {IG12, IG22}
  const D2 *spack;  // 1 per point: {sqrtG, 1/sqrtG}
  int phis_zero;    // phi_s is +0.0 at every point (not read; adding +0.0 is exact)
  const double *f;           // per panel (f differs between panels 5 and 6)
  const double *phi_s;
  const double *p2sp;        // (A*IG) per point, 4 doubles (m11, m21, m12, m22)
  // Line-major copies for the boundary pass (4 doubles per point):
  const double *p2sp_band;   // (A*IG) on the band lines, see band_point()
  const double *sp2p_ghost;  // (G*IA) at the ghost points, see ghost_point()
  const double *sp2p_edge;   // (G*IA) on the edge lines, see band_point() with d = 1
  const double *gcoef;       // (H, nPV): ghost offsets (ghost_mod.f90:11-18)
  const int32_t *giref;      // (H, nPV): 1-based cells
};

// One RK stage of one instance: where the stage reads and writes.
// A time level is stored as {phi, u} pairs and v, so that a point's state
// comes in two loads (pu[g] = {phi, u}, v[g]), reference layout otherwise.
struct StageArgs {
  const D2 *in_pu;                     // stage state Y_s (halo filled)
  const double *in_v;
  const D2 *b_pu;                      // RK base state Y_n
  const double *b_v;
  D2 *acc_pu;                          // RK4 tendency accumulator: {phiG, u} per point
  double *acc_v;                       //   and v
  D2 *o_pu;                            // stage output (interior, raw); the tendency seam
  double *o_v;                         //   writes {phiG, u} and v tendencies on (nPV, nPV, 6)
  double *band_raw;                    // raw copy of the lines nearest each side, see band_slot()
  const double *cu0, *cv0;             // host contravariant winds (first stage only)
  double c1;                           // stage increment coefficient (x dt)
};

struct BoundaryArgs {
  D2 *pu;                              // level to finish (halo + edges)
  double *v;
  const double *band_raw;
};

// Stage kinds of time_scheme_mod.F90:13-123.
enum StageKind : int {
  kTend = 0,   // spatial operator only (tendency seam)
  kRK4S1 = 1,  // Y2 = Yn + dt/2 k1;      acc = k1
  kRK4S2 = 2,  // Y3 = Yn + dt/2 k2;      acc += 2 k2
  kRK4S3 = 3,  // Y4 = Yn + dt k3;        acc += 2 k3
  kRK4S4 = 4,  // Yn+1 = Yn + dt (acc + k4)/6
  kRK3S1 = 5,  // Y1 = Yn + dt k1
  kRK3S2 = 6,  // Y2 = 3/4 Yn + 1/4 Y1 + 1/4 dt k2
  kRK3S3 = 7,  // Yn+1 = Yn/3 + 2/3 Y2 + 2/3 dt k3
};

// The band: the H+1 lines nearest each side (0 W, 1 E, 2 S, 3 N) of each
// panel p (1..6), as the stage kernel wrote them (before the boundary pass).
// Depth d = 1 is the edge line; the along-line index al (1..nPV) is j on the
// W/E sides and i on the S/N sides.  The boundary pass reads only the band,
// contiguously along the lines.
__host__ __device__ inline uint32_t band_point(int npv, int H, int p, int side, int d, int al) {
  return (((uint32_t)(p - 1) * 4u + (uint32_t)side) * (uint32_t)(H + 1) + (uint32_t)(d - 1)) *
             (uint32_t)npv + (uint32_t)(al - 1);
}
// field 0 phi, 1 u, 2 v
__host__ __device__ inline uint32_t band_slot(int npv, int H, int p, int side, int d, int field,
                                              int al) {
  return ((((uint32_t)(p - 1) * 4u + (uint32_t)side) * (uint32_t)(H + 1) + (uint32_t)(d - 1)) * 3u +
          (uint32_t)field) * (uint32_t)npv + (uint32_t)(al - 1);
}
// Ghost point at depth k (1..H) beyond side `side` of panel p, along index s.
__host__ __device__ inline uint32_t ghost_point(int npv, int H, int p, int side, int k, int s) {
  return (((uint32_t)(p - 1) * 4u + (uint32_t)side) * (uint32_t)H + (uint32_t)(k - 1)) *
             (uint32_t)npv + (uint32_t)(s - 1);
}
// (i, j) of band point (side, depth d, along index al).
__host__ __device__ inline void band_ij(int npv, int side, int d, int al, int &i, int &j) {
  if (side == 0) { i = d; j = al; }
  else if (side == 1) { i = npv + 1 - d; j = al; }
  else if (side == 2) { i = al; j = d; }
  else { i = al; j = npv + 1 - d; }
}
// Edge-line point (depth 1) of side `side` of panel p.
__host__ __device__ inline uint32_t edge_point(int npv, int p, int side, int al) {
  return ((uint32_t)(p - 1) * 4u + (uint32_t)side) * (uint32_t)npv + (uint32_t)(al - 1);
}
// Band lines per solve (all panels, sides, depths and fields).
__host__ __device__ inline size_t band_values(int dof, int npv) {
  return (size_t)6 * 4 * dof * 3 * npv;
}

// Work items of the boundary pass on one level: every ghost point (6 panels
// x 4 sides x H depths x nPV), every point of the 12 cube edges (its two
// copies averaged; the end points are skipped), and the 8 cube vertices.
__host__ __device__ inline int boundary_items(int dof, int npv) {
  return 6 * 4 * (dof - 1) * npv + 12 * npv + 8;
}

// (i, j) of point s (1..nPV) along side `side` of a panel.
__host__ __device__ inline void edge_ij(int npv, int side, int s, int &i, int &j) {
  if (side == 0) { i = 1; j = s; }
  else if (side == 1) { i = npv; j = s; }
  else if (side == 2) { i = s; j = 1; }
  else { i = s; j = npv; }
}

// Index arithmetic at full rate.  The 32-bit integer multiply and the
// multiply-high behind a division by a constant issue at a quarter of the
// VALU rate on gfx950; the 24-bit multiply issues at the full rate.
// mul24: a * b for a, b < 2^24 whose product is below 2^32.
__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) { return __umul24(a, b); }
// div_small<D>(t) = t / D for 0 <= t < 2048 and D <= 1024: with
// M = ceil(2^21 / D) = (2^21 + e) / D, 0 <= e < D,
// t * M / 2^21 = t / D + t * e / (D * 2^21), whose floor is floor(t / D)
// while t * e < 2^21 (t < 2048 <= 2^21 / D); t * M < 2048 * (2^21 / 3 + 1)
// stays below 2^32 (D = 3 is the smallest divisor that is not a power of 2).
template <int D>
__device__ __forceinline__ int div_small(int t) {
  static_assert(D >= 1 && D <= 1024, "divisor range");
  if constexpr ((D & (D - 1)) == 0) {
    return (int)((uint32_t)t / (uint32_t)D);
  } else {
    constexpr uint32_t M = ((1u << 21) + D - 1) / D;
    return (int)(mul24((uint32_t)t, M) >> 21);
  }
}

// Element idx of a device array addressed as a uniform base plus a 32-bit
// byte offset, which the global load/store instructions take directly (no
// 64-bit address arithmetic per access).  The library keeps every array that
// the stage kernels touch below 2^32 bytes (n <= 2048).
template <typename T>
__device__ __forceinline__ T ldg(const T *base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      (size_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void stg(T *base, uint32_t idx, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (size_t)(idx * (uint32_t)sizeof(T))) = v;
}

// The same at a byte offset formed by the caller (for 12-byte records, whose
// offset idx * 12 would otherwise take a quarter-rate multiply).
template <typename T>
__device__ __forceinline__ T ldgb(const T *base, uint32_t byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (size_t)byte_off);
}
template <typename T>
__device__ __forceinline__ void stgb(T *base, uint32_t byte_off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (size_t)byte_off) = v;
}

// Linear index of point (i, j) of panel p (1-based, halo-relative i, j).
__host__ __device__ inline long long pidx(int nh, int H, int p, int i, int j) {
  return ((long long)(p - 1) * nh + (j + H - 1)) * (long long)nh + (i + H - 1);
}

// ---------------------------------------------------------------------------
// Cubed-sphere connectivity (ghost_mod.f90:88-152, spatial_operators_mod.F90:
// 298-386).  Sides: 0 W (i = 1), 1 E (i = nPV), 2 S (j = 1), 3 N (j = nPV).
// ---------------------------------------------------------------------------

// Source of the halo of (panel, side) at depth k, along index s: interior
// point of panel q on the line at normal index (far ? nPV-k : k+1), which is
// a column (i fixed) if col, traversed reversed (nPV+1-s) if rev.
struct HaloSrc { int8_t q, col, far, rev; };
// A cube edge shared by (pa, sa) and (pb, sb); b traversed reversed if rev.
struct EdgePair { int8_t pa, sa, pb, sb, rev; };
// A panel corner: panel, i-corner, j-corner (0 -> index 1, 1 -> nPV).
struct VtxPt { int8_t p, ci, cj; };

// One copy per translation unit (no relocatable device code).
// Connectivity tables (ghost_mod.f90:104-146, spatial_operators_mod.F90:298-386).
static __constant__ HaloSrc kHalo[6][4] = {
    /* P1 */ {{4, 1, 1, 0}, {2, 1, 0, 0}, {6, 0, 1, 0}, {5, 0, 0, 0}},
    /* P2 */ {{1, 1, 1, 0}, {3, 1, 0, 0}, {6, 1, 1, 1}, {5, 1, 1, 0}},
    /* P3 */ {{2, 1, 1, 0}, {4, 1, 0, 0}, {6, 0, 0, 1}, {5, 0, 1, 1}},
    /* P4 */ {{3, 1, 1, 0}, {1, 1, 0, 0}, {6, 1, 0, 0}, {5, 1, 0, 1}},
    /* P5 */ {{4, 0, 1, 1}, {2, 0, 1, 0}, {1, 0, 1, 0}, {3, 0, 1, 1}},
    /* P6 */ {{4, 0, 0, 0}, {2, 0, 0, 1}, {3, 0, 0, 1}, {1, 0, 0, 0}},
};
static __constant__ EdgePair kEdges[12] = {
    {1, 0, 4, 1, 0}, {2, 0, 1, 1, 0}, {3, 0, 2, 1, 0}, {4, 0, 3, 1, 0},
    {1, 3, 5, 2, 0}, {1, 2, 6, 3, 0}, {2, 3, 5, 1, 0}, {2, 2, 6, 1, 1},
    {3, 3, 5, 3, 1}, {3, 2, 6, 2, 1}, {4, 3, 5, 0, 1}, {4, 2, 6, 0, 0},
};
static __constant__ VtxPt kVerts[8][3] = {
    {{1, 0, 1}, {4, 1, 1}, {5, 0, 0}}, {{1, 1, 1}, {2, 0, 1}, {5, 1, 0}},
    {{1, 0, 0}, {4, 1, 0}, {6, 0, 1}}, {{1, 1, 0}, {2, 0, 0}, {6, 1, 1}},
    {{3, 0, 1}, {2, 1, 1}, {5, 1, 1}}, {{3, 1, 1}, {4, 0, 1}, {5, 0, 1}},
    {{2, 1, 0}, {3, 0, 0}, {6, 1, 0}}, {{3, 1, 0}, {4, 0, 0}, {6, 0, 0}},
};


// ---------------------------------------------------------------------------
// MCV formulas (FP64, reference evaluation order)
// ---------------------------------------------------------------------------

// polyfitL: one-sided derivative at the cell's left end (spatial_operators_mod.F90:204-218)
template <int DOF>
__device__ __forceinline__ double fit_left(const double *pv, const MCVDivisors &dv) {
  if constexpr (DOF == 3) {
    return -div_by(pv[2] - 4.0 * pv[1] + 3.0 * pv[0], dv.dx);
  } else {
    return div_by(-(11.0 * pv[0]) + 18.0 * pv[1] - 9.0 * pv[2] + 2.0 * pv[3], dv.dx2);
  }
}

// polyfitR: one-sided derivative at the cell's right end (spatial_operators_mod.F90:220-234)
template <int DOF>
__device__ __forceinline__ double fit_right(const double *pv, const MCVDivisors &dv) {
  if constexpr (DOF == 3) {
    return div_by(pv[0] - 4.0 * pv[1] + 3.0 * pv[2], dv.dx);
  } else {
    return div_by(11.0 * pv[3] - 18.0 * pv[2] + 9.0 * pv[1] - 2.0 * pv[0], dv.dx2);
  }
}

// Rusanov flux on derivatives (spatial_operators_mod.F90:236-246)
__device__ __forceinline__ double riemann(double fxL, double fxR, double qxL, double qxR,
                                          double lam) {
  return 0.5 * ((fxL + fxR) - lam * (qxR - qxL));
}

// Largest gravity-wave speed along one direction (spatial_operators_mod.F90:248-280)
__device__ __forceinline__ double wave_speed(double c, double phi, double gii) {
  double l1 = c - sqrt(gii * phi), l2 = c + sqrt(gii * phi);
  // the larger of two non-negative values (both NaN together or neither)
  return fmax(fabs(l1), fabs(l2));
}

// c/|c| with 0 -> +1 (spatial_operators_mod.F90:83-85): exactly +-1 for every
// finite c, NaN for NaN.
__device__ __forceinline__ double upwind_sign(double c) {
  if (c != c) return c;
  return c == 0.0 ? 1.0 : copysign(1.0, c);
}

// Interior-point derivative of a cell from its point values and the two
// cell-end derivatives (spatial_operators_mod.F90:178-200).  pos in [1, H-1].
template <int DOF>
__device__ __forceinline__ double cell_interior(const double *f, double dL, double dR,
                                                int pos, double dx, const MCVDivisors &dv) {
  if constexpr (DOF == 3) {
    double via = div_by(f[2] - f[0], dv.dx);
    return 1.5 * via - 0.25 * (dL + dR);
  } else {
    // pos 1: 4/3 via - (4 dL + 5 dR)/27 - 4/27 dx fxx
    // pos 2: 4/3 via - (5 dL + 4 dR)/27 + 4/27 dx fxx
    // written with selects so lanes at either position take the same path
    constexpr Recip k27{27.0, 1.0 / 27.0};
    const double via = div_by(f[3] - f[0], dv.dx);
    const double fxx = div_by(4.5 * (f[0] - f[1] - f[2] + f[3]), dv.dxx);
    const bool first = pos == 1;
    const double cl = first ? 4.0 : 5.0, cr = first ? 5.0 : 4.0;
    const double a = 4.0 / 3.0 * via - div_by(cl * dL + cr * dR, k27);
    const double c = 4.0 / 27.0 * dx * fxx;
    return first ? a - c : a + c;
  }
}

// Degree DOF-1 Lagrange interpolation of a cell at offset c (ghost_mod.f90:183-227)
template <int DOF>
__device__ __forceinline__ double ghost_interp(const double *q, double c, double dx) {
  if constexpr (DOF == 3) {
    return q[0] - (3.0 * q[0] - 4.0 * q[1] + q[2]) * c / dx +
           2.0 * (q[0] - 2.0 * q[1] + q[2]) * (c * c) / (dx * dx);
  } else {
    double dx3 = dx * dx * dx, c3 = c * c * c;
    return (2.0 * dx3 * q[0] +
            (dx * dx) * (-(11.0 * q[0]) + 18.0 * q[1] - 9.0 * q[2] + 2.0 * q[3]) * c +
            9.0 * dx * (2.0 * q[0] - 5.0 * q[1] + 4.0 * q[2] - q[3]) * (c * c) +
            9.0 * (-q[0] + 3.0 * q[1] - 3.0 * q[2] + q[3]) * c3) /
           (2.0 * dx3);
  }
}

// 2x2 apply with M = (m11, m21, m12, m22) (projection_mod.F90:60-96)
__device__ __forceinline__ void apply2(const double *M, double a, double b, double &o1,
                                       double &o2) {
  o1 = M[0] * a + M[2] * b;
  o2 = M[1] * a + M[3] * b;
}

// blockIdx -> logical work index so that the blocks one XCD runs are
// consecutive in the logical order.  Blocks are dealt round-robin to the 8
// XCDs of an MI355X, each with its own L2; consecutive logical indices touch
// neighbouring tiles and the same metric tile on the six panels.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  constexpr int kXcd = 8;
  int q = nb / kXcd, r = nb % kXcd;
  int x = b % kXcd, slot = b / kXcd;
  return x * q + (x < r ? x : r) + slot;
}

}  // namespace mcv
