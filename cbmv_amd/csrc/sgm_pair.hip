// Launchers of the paired SGM sweeps (sgm_pair.hpp) for the fused pipeline.
#include "sgm_pair.hpp"

namespace cbmv {

namespace {

int pair_dpl(int D) { return D <= 64 ? 1 : D <= 128 ? 2 : 4; }

template <int DPL, int PF, bool ROWS>
void pair_go(const float* x0, const float* x1p, const float* in, float* out, int H, int W, int D,
             const Penalties& pen, float tau_so, int direction, int sa, int sd, int mode,
             double scale, PairSync sync, long long spin, hipStream_t s) {
  const int L = ROWS ? H : W;
  constexpr int K = pair_waves<ROWS>();
  const dim3 g((L + K - 1) / K), b(64 * K);
  const bool full = D == 64 * DPL;
#define CBMV_PAIR_GO(DR, F, LA, FU)                                                          \
  k_sgm_pair<DPL, PF, DR, F, LA, FU, ROWS><<<g, b, 0, s>>>(x0, x1p, in, out, H, W, D, pen, \
                                                           tau_so, sa, sd, sync, scale, spin)
#define CBMV_PAIR_MODES(DR, FU)                         \
  switch (mode & 3) {                                   \
    case 0: CBMV_PAIR_GO(DR, false, false, FU); break;  \
    case 1: CBMV_PAIR_GO(DR, true, false, FU); break;   \
    case 2: CBMV_PAIR_GO(DR, false, true, FU); break;   \
    default: CBMV_PAIR_GO(DR, true, true, FU); break;   \
  }
  if (direction > 0) {
    if (full) {
      CBMV_PAIR_MODES(1, true)
    } else {
      CBMV_PAIR_MODES(1, false)
    }
  } else {
    if (full) {
      CBMV_PAIR_MODES(-1, true)
    } else {
      CBMV_PAIR_MODES(-1, false)
    }
  }
#undef CBMV_PAIR_MODES
#undef CBMV_PAIR_GO
}

// PFR, PFC: steps of prefetch of the row and the column sweeps
template <int DPL, int PFR, int PFC>
void pair_dispatch(const float* x0, const float* x1p, const float* in, float* out, int H, int W,
                   int D, const Penalties& pen, float tau_so, int direction, int pair, int mode,
                   double scale, PairSync sync, long long spin, hipStream_t s) {
  // pair 0: E + SE, 1: W + NW (scanlines are rows); 2: S + SW, 3: N + NE (columns)
  static const int ksa[4] = {1, -1, 1, -1}, ksd[4] = {1, -1, -1, 1};
  if (pair < 2)
    pair_go<DPL, PFR, true>(x0, x1p, in, out, H, W, D, pen, tau_so, direction, ksa[pair],
                           ksd[pair], mode, scale, sync, spin, s);
  else
    pair_go<DPL, PFC, false>(x0, x1p, in, out, H, W, D, pen, tau_so, direction, ksa[pair],
                            ksd[pair], mode, scale, sync, spin, s);
}

}  // namespace

int64_t sgm_pair_edge_floats(int H, int W, int D) {
  const int64_t ndp = 64 * pair_dpl(D);
  constexpr int KR = pair_waves<true>(), KC = pair_waves<false>();
  const int64_t r = (int64_t)((H + KR - 1) / KR - 1) * W;
  const int64_t c = (int64_t)((W + KC - 1) / KC - 1) * H;
  return std::max<int64_t>(std::max(r, c), 1) * ndp;
}

int64_t sgm_pair_edge_entries(int H, int W, int D) {
  return sgm_pair_edge_floats(H, W, D) / (64 * pair_dpl(D));
}

int launch_sgm_pair_fill(float* edge, int64_t n_edge, float* edge_mn, int64_t n_mn, unsigned* ctr,
                         bool always, hipStream_t s) {
  PairSync sync{edge, edge_mn, ctr};
  k_pair_fill<<<always ? 1024 : 64, 256, 0, s>>>(sync, n_edge, n_mn, always);
  k_pair_repaired<<<1, 64, 0, s>>>(sync);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_sgm_pair_f32(const float* x0, const float* x1p, const float* in, float* out, int H,
                        int W, int D, float pi1, float pi2, float tau_so, float alpha1, float q1,
                        float q2, int direction, int pair, int mode, double scale, float* edge,
                        float* edge_mn, unsigned* ctr, long long spin, hipStream_t s) {
  CBMV_REQUIRE(D >= 1 && D <= 256, "paired SGM sweeps support 1 <= ndisp <= 256 (got %d)", D);
  CBMV_REQUIRE(pair >= 0 && pair < 4, "bad SGM path pair %d", pair);
  CBMV_REQUIRE(direction == 1 || direction == -1, "SGM direction must be +1 or -1 (got %d)",
               direction);
  const Penalties pen = make_penalties(pi1, pi2, alpha1, q1, q2);
  for (int c = 0; c < 3; c++)
    CBMV_REQUIRE(pen.p1[c] >= 0.f && pen.p2[c] >= 0.f && pen.p1a[c] >= 0.f,
                 "SGM on float volumes needs non-negative penalties (pi1, pi2, q1, q2, alpha1)");
  PairSync sync{edge, edge_mn, ctr};
  switch (pair_dpl(D)) {
    case 1: pair_dispatch<1, 6, 6>(x0, x1p, in, out, H, W, D, pen, tau_so, direction, pair, mode, scale, sync, spin, s); break;
    case 2: pair_dispatch<2, 6, 4>(x0, x1p, in, out, H, W, D, pen, tau_so, direction, pair, mode, scale, sync, spin, s); break;
    default: pair_dispatch<4, 4, 3>(x0, x1p, in, out, H, W, D, pen, tau_so, direction, pair, mode, scale, sync, spin, s); break;
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

}  // namespace cbmv
