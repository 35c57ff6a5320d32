// Launchers of the chirp-z kernels for one precision, dispatched on the
// convolution length L (included by czt_f32.hip / czt_f64.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "czt_api.h"
#include "czt_kernels.h"

namespace mjx {

template <typename T>
CztPlan<T> czt_dev(const CztTables& h) {
  const cpx<T>* t;
  if constexpr (std::is_same<T, float>::value) t = h.t32;
  else t = h.t64;
  CztPlan<T> p;
  p.c = t;
  p.d = t + h.n;
  p.e = t + 2 * h.n;
  p.n = h.n;
  return p;
}

inline int czt_grid_check(long long blocks) {
  if (blocks < 1 || blocks > 0x7fffffffLL) return fail(MJX_E_LIMIT, "chirp-z launch too large");
  return MJX_OK;
}

// Calls F<L>::run(args...) for the runtime convolution length.
template <template <int> class F, typename... A>
int czt_with_len(int L, A... args) {
#define X(LL) \
  if (L == LL) return F<LL>::run(args...);
  MJX_FOR_CZT_LENGTHS(X)
#undef X
  return fail(MJX_E_SIZE, "no chirp-z kernel for convolution length " + std::to_string(L));
}

template <typename T>
struct CztLaunch {
  template <int L>
  struct FwdPack {
    static int run(int B, int ny, int nh, const T* a, const T* b, long long pitch, long long istride, cpx<T>* W,
                   T in_scale, CztPlan<T> p, hipStream_t st) {
      using C = CztCfg<T, L>;
      const int tiles = (ny + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)tiles * B);
      if (rc) return rc;
      hipLaunchKernelGGL((k_czt_row_fwd_pack<T, L>), dim3(tiles * B), dim3(C::BT), 0, st, a, b, pitch, istride, W,
                         ny, nh, tiles, in_scale, p, twiddles<T>());
      return check_launch("czt_row_fwd_pack");
    }
  };
  template <int L>
  struct InvUnpack {
    static int run(int B, int ny, int nh, const cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                   int ghosts, T keep, const InstParams* react, CztPlan<T> p, hipStream_t st) {
      using C = CztCfg<T, L>;
      const int tiles = (ny + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)tiles * B);
      if (rc) return rc;
      hipLaunchKernelGGL((k_czt_row_inv_unpack<T, L>), dim3(tiles * B), dim3(C::BT), 0, st, W, a, b, pitch, istride,
                         ny, nh, tiles, ghosts, keep, react, p, twiddles<T>());
      return check_launch("czt_row_inv_unpack");
    }
  };
  template <int L>
  struct RowGs {
    static int run(int B, int ny, int nh, cpx<T>* W, const InstParams* ip, int rxn, int mstep, CztPlan<T> p,
                   hipStream_t st) {
      using C = CztCfg<T, L>;
      const int tiles = (ny + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)tiles * B);
      if (rc) return rc;
      hipLaunchKernelGGL((k_czt_row_gs<T, L>), dim3(tiles * B), dim3(C::BT), 0, st, W, ip, ny, nh, tiles, rxn, mstep,
                         p, twiddles<T>());
      return check_launch("czt_row_gs");
    }
  };
  template <int L>
  struct ColApply {
    static int run(int B, int nx, int nh, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                   const InstParams* ip, int nrep, T s, int mstep, CztPlan<T> p, hipStream_t st) {
      using C = CztCfg<T, L>;
      const int tiles = (2 * nh + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)tiles * B);
      if (rc) return rc;
      hipLaunchKernelGGL((k_czt_col_apply<T, L>), dim3(tiles * B), dim3(C::BT), 0, st, W, nx, nh, tiles, lamx, lamy,
                         lam2d, ip, nrep, s, mstep, p, twiddles<T>());
      return check_launch("czt_col_apply");
    }
  };
  template <int L>
  struct R2R {
    static int run(CztR2RArgs a, int orient, bool apply, CztPlan<T> p1, CztPlan<T> p2, int B, hipStream_t st) {
      using C = CztCfg<T, L>;
      a.tiles = (a.nlines + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)a.tiles * B);
      if (rc) return rc;
      const dim3 g(a.tiles * B), b(C::BT);
      const cpx<T>* tw = twiddles<T>();
      if (orient == CZT_ROW) {
        if (apply) hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_ROW, true>), g, b, 0, st, a, p1, p2, tw);
        else hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_ROW, false>), g, b, 0, st, a, p1, p2, tw);
      } else {
        if (apply) hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_COL, true>), g, b, 0, st, a, p1, p2, tw);
        else hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_COL, false>), g, b, 0, st, a, p1, p2, tw);
      }
      return check_launch("czt_r2r");
    }
  };

  static int row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride,
                          cpx<T>* W, hipStream_t st, double in_scale) {
    CztTables h;
    int rc = czt_tables(nx, nx, 0, 0, h);
    if (rc) return rc;
    return czt_with_len<FwdPack>(h.L, B, ny, nx / 2 + 1, a, b, pitch, istride, W, (T)in_scale, czt_dev<T>(h), st);
  }
  static int row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch,
                            long long istride, int ghosts, hipStream_t st, double keep, const InstParams* react) {
    CztTables h;
    int rc = czt_tables(nx, nx, 0, 0, h);
    if (rc) return rc;
    return czt_with_len<InvUnpack>(h.L, B, ny, nx / 2 + 1, W, a, b, pitch, istride, ghosts, (T)keep, react,
                                   czt_dev<T>(h), st);
  }
  static int row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep) {
    CztTables h;
    int rc = czt_tables(nx, nx, 0, 0, h);
    if (rc) return rc;
    return czt_with_len<RowGs>(h.L, B, ny, nx / 2 + 1, W, ip, rxn, mstep, czt_dev<T>(h), st);
  }
  static int col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                       const InstParams* ip, int nrep, hipStream_t st, int mstep) {
    CztTables h;
    int rc = czt_tables(ny, ny, 0, 0, h);
    if (rc) return rc;
    const T s = (T)(1.0 / ((double)nx * (double)ny));
    return czt_with_len<ColApply>(h.L, B, nx, nx / 2 + 1, W, lamx, lamy, lam2d, ip, nrep, s, mstep, czt_dev<T>(h),
                                  st);
  }
  static int r2r(const CztR2RArgs& a, int orient, bool apply, const CztTables& h1, const CztTables& h2, int B,
                 hipStream_t st) {
    return czt_with_len<R2R>(h1.L, a, orient, apply, czt_dev<T>(h1), czt_dev<T>(h2), B, st);
  }
};

}  // namespace mjx
