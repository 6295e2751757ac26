// One SGM path over every scanline of an [H,W,D] volume (post.cu:178-259,
// generalised to the two diagonal directions), shared by the float32 and
// float64 launchers (sgm_f32.hip, sgm_f64.hip; one translation unit each so
// that they compile in parallel).
//
// One wavefront per scanline (four scanlines per workgroup, no barriers).
// Lane l owns the DPL consecutive disparities d = l*DPL + j, so the cost and
// accumulator rows of a pixel (contiguous in d) are one coalesced load per
// lane; d-1 / d+1 of the lane's edge disparities come from the neighbouring
// lanes by one DPP move each, and the minimum over d is a DPP wave reduction.
// The inputs of the next PF - 1 steps (cost, accumulator, image values) are
// loaded ahead into a ring of PF register slots of raw values: slot k holds
// the steps s = k (mod PF), and step s refills slot s - 1 (whose values it
// reads as the previous pixel's), so no loaded value is ever copied between
// registers: a copy would wait for the load, and the in-order vmcnt counter
// would make it wait for every later load too.
//
// With one wavefront per SIMD every instruction of the step costs an issue
// slot, so the step keeps the per-disparity work small: everything that is
// the same for the whole scanline (positions, pointers, the D1 test and the
// penalty class it selects) runs on scalar registers, and each disparity
// does two compares and three selects for its penalties.
//
// Code-generation notes: the step lambda captures few variables (the
// penalty table is one VGPR, value i in lane i, read with v_readlane): with
// around 35 captures its closure was not split into registers and every
// captured array went to scratch memory; and a scheduling barrier opens each
// step, since computing on ring values early makes the step wait for loads
// issued in the last few steps.
#pragma once

#include <cmath>

#include "ops.hpp"

namespace cbmv {
namespace {

struct Penalties {
  float p1[3], p2[3], p1a[3];  // [0] both < tau, [1] both > tau, [2] otherwise; p1a = p1/alpha1
};

Penalties make_penalties(float pi1, float pi2, float alpha1, float q1, float q2) {
  // float arithmetic of post.cu:239-255, evaluated on the host in IEEE single
  Penalties p;
  volatile float q12 = q1 * q2;
  p.p1[0] = pi1;
  p.p2[0] = pi2;
  p.p1[1] = pi1 / q12;
  p.p2[1] = pi2 / q12;
  p.p1[2] = pi1 / q1;
  p.p2[2] = pi2 / q1;
  for (int i = 0; i < 3; i++) p.p1a[i] = p.p1[i] / alpha1;
  return p;
}

void path_dir(int path, int& dx, int& dy) {
  static const int kdx[8] = {1, -1, 0, 0, 1, -1, 1, -1};
  static const int kdy[8] = {0, 0, 1, -1, 1, -1, -1, 1};
  dx = kdx[path];
  dy = kdy[path];
}

// Contiguous per-lane vector of N elements (one global_load_dwordxK when the
// address allows; `p` need only be element-aligned).
template <typename T, int N>
__device__ __forceinline__ void vload(const T* p, T (&v)[N]) {
  struct alignas(sizeof(T)) Chunk {
    T e[N];
  } c;
  __builtin_memcpy(&c, p, sizeof(c));
#pragma unroll
  for (int j = 0; j < N; j++) v[j] = c.e[j];
}

template <typename T, int N>
__device__ __forceinline__ void vstore(T* p, const T (&v)[N]) {
  struct alignas(sizeof(T)) Chunk {
    T e[N];
  } c;
#pragma unroll
  for (int j = 0; j < N; j++) c.e[j] = v[j];
  __builtin_memcpy(p, &c, sizeof(c));
}

// A copy of a value in a VGPR, so that it can sit next to an SGPR operand or
// a lane mask in one VALU instruction (a gfx9 VALU instruction reads a single
// scalar value).
__device__ __forceinline__ float in_vgpr(float v) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// A zero the compiler cannot see through: added to a uniform address, it
// keeps a load on the vector memory path (in-order vmcnt, like the rest of
// the ring) instead of the scalar cache.
__device__ __forceinline__ int opaque_zero() {
  int r;
  asm("v_mov_b32 %0, 0" : "=v"(r));
  return r;
}

template <typename T, int DPL, int PF, int DIR, bool FIRST, bool LAST, bool FULL>
__global__ void __launch_bounds__(256) k_sgm_wave(const float* __restrict__ x0,
                                                  const float* __restrict__ x1p,
                                                  const T* __restrict__ in, T* __restrict__ out,
                                                  int H, int W, int D, int64_t sy, int64_t sx,
                                                  Penalties pen, float tau_so, int dx, int dy,
                                                  int nlines, double scale) {
  const int lane = threadIdx.x & 63;
  // wave-uniform (readfirstlane): positions along the scanline and the
  // offsets below live in scalar registers; lanes differ only by their
  // disparity offset
  const int line = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (line >= nlines) return;  // whole wavefront; no block-level synchronisation below
  int xs, ys;
  if (dy == 0) {
    ys = line;
    xs = dx > 0 ? 0 : W - 1;
  } else if (dx == 0) {
    xs = line;
    ys = dy > 0 ? 0 : H - 1;
  } else if (line < W) {
    xs = line;
    ys = dy > 0 ? 0 : H - 1;
  } else {
    xs = dx > 0 ? 0 : W - 1;
    ys = (dy > 0) ? (line - W + 1) : (line - W);
  }
  const int lx = dx > 0 ? W - xs : (dx < 0 ? xs + 1 : 0x7fffffff);
  const int ly = dy > 0 ? H - ys : (dy < 0 ? ys + 1 : 0x7fffffff);
  const int len = min(lx, ly);
  const int dbase = lane * DPL;
  const int64_t pitch = (int64_t)W + 2 * kSgmPad;
  int joff[DPL];
  bool jok[DPL];
  // the right-image pair (x + d DIR, x + d DIR - dx) of post.cu:229-234 is
  // inside the image iff xlo[j] <= x < xlo[j] + xw[j]
  // (one unsigned compare: x - xlo[j] < xw[j])
  int xlo[DPL];
  uint32_t xw[DPL];
#pragma unroll
  for (int j = 0; j < DPL; j++) {
    jok[j] = FULL || dbase + j < D;
    joff[j] = FULL ? dbase + j : min(dbase + j, D - 1);
    const int dd = (dbase + j) * DIR;
    xlo[j] = max(0, dx) - dd;
    xw[j] = (uint32_t)max(W + min(0, dx) - dd - xlo[j], 0);
  }
  // Uniform byte offsets of the next step to fetch and of the next store,
  // with their per-step increments; lanes add a constant 32-bit offset. The
  // lane's right-image segment starts at column x + seg0; for DIR < 0 the
  // uniform part is shifted by 64 DPL columns so that the lane part is
  // non-negative.
  const int seg0 = DIR > 0 ? dbase : -(dbase + DPL - 1);
  constexpr int kShift = DIR > 0 ? 0 : 64 * DPL;
  const int64_t vstep = (dy * sy + dx * sx) * (int64_t)sizeof(T);
  const int64_t istep = ((int64_t)dy * W + dx) * 4, pstep = (dy * pitch + dx) * 4;
  const uint32_t lane_v = (uint32_t)(dbase * sizeof(T));
  const uint32_t lane_p = (uint32_t)((seg0 + kShift) * 4);
  const int lane_0 = opaque_zero();
  // running byte pointers (uniform) of the next step to fetch and the next store
  const char* pin = (const char*)in + (ys * sy + xs * sx) * (int64_t)sizeof(T);
  const char* pacc = (const char*)out + (ys * sy + xs * sx) * (int64_t)sizeof(T);
  const char* px1 = (const char*)x1p + (ys * pitch + kSgmPad + xs - kShift) * 4;
  const char* px0 = (const char*)x0 + ((int64_t)ys * W + xs) * 4;
  char* pst = (char*)out + (ys * sy + xs * sx) * (int64_t)sizeof(T);

  // ring of raw loaded values; nothing is computed from them before use
  T rin[PF][DPL], rout[PF][DPL];
  float rseg[PF][DPL], rx0[PF];

  auto fetch = [&](int slot) {
    if (FULL) {
      vload((const T*)(pin + lane_v), rin[slot]);
      if (!FIRST) vload((const T*)(pacc + lane_v), rout[slot]);
    } else {
#pragma unroll
      for (int j = 0; j < DPL; j++) {
        rin[slot][j] = ((const T*)pin)[joff[j]];
        if (!FIRST) rout[slot][j] = ((const T*)pacc)[joff[j]];
      }
    }
    vload((const float*)(px1 + lane_p), rseg[slot]);
    rx0[slot] = *(const float*)(px0 + lane_0);
  };
  // advance the fetch pointers unless the line's last step has been fetched
  // (later fetches then repeat it)
  auto advance = [&](bool more) {
    if (more) {
      pin += vstep;
      pacc += vstep;
      px0 += istep;
      px1 += pstep;
    }
  };

  // steps 0 .. PF-2 into slots 0 .. PF-2; slot PF-1 ("step -1") holds the
  // previous-pixel values of step 0: zero segment and image value
#pragma unroll
  for (int k = 0; k < PF - 1; k++) {
    fetch(k);
    advance(k + 1 < len);
  }
#pragma unroll
  for (int j = 0; j < DPL; j++) rseg[PF - 1][j] = 0.f;
  rx0[PF - 1] = 0.f;

  // Penalties for the d-1 (left) and d+1 (right) neighbours and P2, per
  // class: the d-1 term uses P1/alpha1 on downward paths, the d+1 term on
  // upward ones (post.cu:252-255). Scalar registers; the mixed class [2]
  // also as VGPRs for the per-disparity selects.
  const bool down = dy > 0, up = dy < 0;
  // The nine values sit in one VGPR, value i in lane i; a step reads its
  // class's three with v_readlane at a uniform index (selects among
  // registers or memory here turn into indexed loads).
  float ptab = 0.f;
  {
    const float v[9] = {down ? pen.p1a[0] : pen.p1[0], down ? pen.p1a[1] : pen.p1[1],
                        down ? pen.p1a[2] : pen.p1[2], up ? pen.p1a[0] : pen.p1[0],
                        up ? pen.p1a[1] : pen.p1[1],   up ? pen.p1a[2] : pen.p1[2],
                        pen.p2[0],                     pen.p2[1],
                        pen.p2[2]};
#pragma unroll
    for (int i = 0; i < 9; i++) ptab = lane == i ? v[i] : ptab;
  }
  auto ptab_at = [&](int i) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ptab), i));
  };
  const float Ml = in_vgpr(ptab_at(2)), Mr = in_vgpr(ptab_at(5)), M2 = in_vgpr(ptab_at(8));
  const float ten = in_vgpr(10.f), inf = in_vgpr(INFINITY);
  // D1 >= 0, so its bits order like signed integers, also against a
  // negative tau (tau = -0 taken as +0)
  const int taui = __float_as_int(tau_so + 0.f);
  float prev[DPL];
#pragma unroll
  for (int j = 0; j < DPL; j++) prev[j] = 0.f;
  float mn = 0.f;
  int x = xs;

  auto step = [&](int k, int s) {
    // nothing of this step is scheduled into the previous one: computing
    // on ring values early would wait for their loads early
    __builtin_amdgcn_sched_barrier(0);
    const int kp = (k + PF - 1) % PF;  // slot of the previous pixel
    // neighbours across lanes (DPP wave shifts; the lane without a source
    // keeps the old value, inf: disparities -1 and D do not exist)
    const float left_nb = __int_as_float(__builtin_amdgcn_update_dpp(
        __float_as_int(inf), __float_as_int(prev[DPL - 1]), 0x138, 0xf, 0xf, false));
    const float right_nb = __int_as_float(__builtin_amdgcn_update_dpp(
        __float_as_int(inf), __float_as_int(prev[0]), 0x130, 0xf, 0xf, false));
    const int d1i = __builtin_amdgcn_readfirstlane(__float_as_int(fabsf(rx0[k] - rx0[kp])));
    const bool lt1 = d1i < taui, gt1 = d1i > taui;
    const int cls = lt1 ? 0 : (gt1 ? 1 : 2);
    // D2's class test as one compare: sg D2 < tb with (sg, tb) = (1, tau)
    // when D1 < tau, (-1, -tau) when D1 > tau, (0, 0) (never true) otherwise
    const float sg = lt1 ? 1.f : (gt1 ? -1.f : 0.f);
    const float tb = lt1 ? tau_so : (gt1 ? -tau_so : 0.f);
    const float vAl = ptab_at(cls), vAr = ptab_at(3 + cls), vA2 = ptab_at(6 + cls);
    const double mnd = (double)mn;
    float cur[DPL];
    double val[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) {
      const int d = dbase + j;
      const float seg = DIR > 0 ? rseg[k][j] : rseg[k][DPL - 1 - j];
      const float pseg = DIR > 0 ? rseg[kp][j] : rseg[kp][DPL - 1 - j];
      const bool ok2 = (uint32_t)(x - xlo[j]) < xw[j];
      const float D2 = ok2 ? fabsf(seg - pseg) : ten;
      // both below / both above tau: the step's class value; else the mixed one
      const bool cd = sg * D2 < tb;
      const float Pl = cd ? vAl : Ml;
      const float Pr = cd ? vAr : Mr;
      const float P2 = cd ? vA2 : M2;
      const float pl = j > 0 ? prev[j - 1] : left_nb;
      float pr = j < DPL - 1 ? prev[j + 1] : right_nb;
      if (j == DPL - 1 && !FULL) pr = d + 1 < D ? pr : inf;
      const float c = fminf(fminf(prev[j], mn + P2), fminf(pl + Pl, pr + Pr));
      if constexpr (sizeof(T) == 4) {
        // first pixel of a scanline: the cost itself (post.cu:211-213). There
        // prev = 0 and mn = 0, so with non-negative penalties (checked by the
        // launcher) c = 0 and the general expression is exact.
        val[j] = ((double)rin[k][j] + (double)c) - mnd;
      } else {
        // float64 volumes: the first pixel's cost rounded to float
        val[j] = s == 0 ? (double)(float)rin[k][j] : ((double)rin[k][j] + (double)c) - mnd;
      }
      cur[j] = (float)val[j];
    }
    float lmin = INFINITY;
    T o[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) {
      double v = (FIRST ? 0.0 : (double)rout[k][j]) + val[j];
      if (LAST) v *= scale;
      o[j] = (T)v;
      if (jok[j]) {
        lmin = fminf(lmin, cur[j]);
      } else {
        cur[j] = INFINITY;
      }
      prev[j] = cur[j];
    }
    if (FULL) {
      vstore((T*)(pst + lane_v), o);
    } else {
#pragma unroll
      for (int j = 0; j < DPL; j++)
        if (jok[j]) ((T*)pst)[joff[j]] = o[j];
    }
    pst += vstep;
    x += dx;
    mn = wave_min_dpp(lmin);
  };

  // main part: every fetch is a step of the line
  int s0 = 0;
  for (; s0 + 2 * PF <= len; s0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      step(k, s0 + k);
      fetch((k + PF - 1) % PF);
      advance(true);
    }
  }
  for (; s0 < len; s0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int s = s0 + k;
      if (s >= len) break;
      step(k, s);
      fetch((k + PF - 1) % PF);
      advance(s + PF < len);
    }
  }
}

template <typename T, int DPL, int PF>
void sgm_launch(const float* x0, const float* x1p, const T* in, T* out, int H, int W, int D,
                VolStrides st, const Penalties& pen, float tau_so, int direction, int dx, int dy,
                int lines, int mode, double scale, hipStream_t s) {
  const dim3 g((lines + 3) / 4), b(256);
  const bool full = D == 64 * DPL;
#define CBMV_SGM_GO(DR, F, L, FU)                                                          \
  k_sgm_wave<T, DPL, PF, DR, F, L, FU><<<g, b, 0, s>>>(x0, x1p, in, out, H, W, D, st.sy, \
                                                       st.sx, pen, tau_so, dx, dy, lines, scale)
#define CBMV_SGM_MODES(DR, FU)                             \
  switch (mode & 3) {                                      \
    case 0: CBMV_SGM_GO(DR, false, false, FU); break;      \
    case 1: CBMV_SGM_GO(DR, true, false, FU); break;       \
    case 2: CBMV_SGM_GO(DR, false, true, FU); break;       \
    default: CBMV_SGM_GO(DR, true, true, FU); break;       \
  }
  if (direction > 0) {
    if (full) {
      CBMV_SGM_MODES(1, true)
    } else {
      CBMV_SGM_MODES(1, false)
    }
  } else {
    if (full) {
      CBMV_SGM_MODES(-1, true)
    } else {
      CBMV_SGM_MODES(-1, false)
    }
  }
#undef CBMV_SGM_MODES
#undef CBMV_SGM_GO
}

template <typename T>
int sgm_path(const float* x0, const float* x1p, const T* in, T* out, int H, int W, int D,
             VolStrides st, float pi1, float pi2, float tau_so, float alpha1, float q1, float q2,
             int direction, int path, int mode, double scale, hipStream_t s) {
  CBMV_REQUIRE(D >= 1 && D <= 512, "SGM supports 1 <= ndisp <= 512 (got %d)", D);
  CBMV_REQUIRE(st.sd == 1, "SGM needs the disparities of a pixel contiguous ([H,W,D])");
  CBMV_REQUIRE(path >= 0 && path < 8, "bad SGM path %d", path);
  CBMV_REQUIRE(direction == 1 || direction == -1, "SGM direction must be +1 or -1 (got %d)",
               direction);
  int dx, dy;
  path_dir(path, dx, dy);
  const int lines = dy == 0 ? H : (dx == 0 ? W : W + H - 1);
  const Penalties pen = make_penalties(pi1, pi2, alpha1, q1, q2);
  constexpr bool f64 = sizeof(T) == 8;
  if (!f64) {
    // the float32 scan takes the first pixel of a line through the general
    // step, exact only when no penalty is negative
    for (int c = 0; c < 3; c++)
      CBMV_REQUIRE(pen.p1[c] >= 0.f && pen.p2[c] >= 0.f && pen.p1a[c] >= 0.f,
                   "SGM on float volumes needs non-negative penalties (pi1, pi2, q1, q2, alpha1)");
  }
  // ring depth PF (PF - 1 steps in flight) per lane width
  if (D <= 64)
    sgm_launch<T, 1, f64 ? 5 : 9>(x0, x1p, in, out, H, W, D, st, pen, tau_so, direction, dx, dy,
                                  lines, mode, scale, s);
  else if (D <= 128)
    sgm_launch<T, 2, f64 ? 5 : 9>(x0, x1p, in, out, H, W, D, st, pen, tau_so, direction, dx, dy,
                                  lines, mode, scale, s);
  else if (D <= 256)
    sgm_launch<T, 4, f64 ? 4 : 7>(x0, x1p, in, out, H, W, D, st, pen, tau_so, direction, dx, dy,
                                  lines, mode, scale, s);
  else
    sgm_launch<T, 8, f64 ? 3 : 4>(x0, x1p, in, out, H, W, D, st, pen, tau_so, direction, dx, dy,
                                  lines, mode, scale, s);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

}  // namespace
}  // namespace cbmv
