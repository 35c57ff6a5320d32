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
      hipLaunchKernelGGL((k_czt_row_gs<T, L, false>), dim3(tiles * B), dim3(C::BT), 0, st, W, W, ip, ny, nh, tiles,
                         rxn, mstep, p, twiddles<T>());
      return check_launch("czt_row_gs");
    }
  };
  template <int L>
  struct RowNl {
    static int run(int B, int ny, int nh, const cpx<T>* Win, cpx<T>* Wout, const InstParams* ip, CztPlan<T> p,
                   hipStream_t st) {
      using C = CztCfg<T, L>;
      const int tiles = (ny + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)tiles * B);
      if (rc) return rc;
      hipLaunchKernelGGL((k_czt_row_gs<T, L, true>), dim3(tiles * B), dim3(C::BT), 0, st, Win, Wout, ip, ny, nh,
                         tiles, 0, -1, p, twiddles<T>());
      return check_launch("czt_row_nl");
    }
  };
  template <int STAGE>
  struct ColEtd {
    template <int L>
    struct F {
      static int run(int B, int nh, int sw, const cpx<T>* Win, const cpx<T>* Win2, cpx<T>* Uh, cpx<T>* Nn,
                     cpx<T>* Fh, const T* coef, cpx<T>* Wout, CztPlan<T> p, hipStream_t st) {
        using C = CztCfg<T, L>;
        const int tiles = (2 * nh + C::NL - 1) / C::NL;
        int rc = czt_grid_check((long long)tiles * B);
        if (rc) return rc;
        hipLaunchKernelGGL((k_czt_col_etd<T, L, STAGE>), dim3(tiles * B), dim3(C::BT), 0, st, Win, Win2, Uh, Nn, Fh,
                           coef, Wout, nh, sw, tiles, p, twiddles<T>());
        return check_launch("czt_col_etd");
      }
    };
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
  // Lines per workgroup: CztCfg's default, or a quarter of it when the
  // default would leave most CUs idle (one 512^2 instance: 128 workgroups);
  // every line is its own transform, so the results are the same bitwise.
  template <int L>
  struct R2R {
    template <int NLW>
    static int go(CztR2RArgs a, int orient, bool apply, CztPlan<T> p1, CztPlan<T> p2, int B, hipStream_t st) {
      using C = CztR2RCfg<T, L, NLW>;
      a.tiles = (a.nlines + C::NL - 1) / C::NL;
      int rc = czt_grid_check((long long)a.tiles * B);
      if (rc) return rc;
      const dim3 g(a.tiles * B), b(C::BT);
      const cpx<T>* tw = twiddles<T>();
      if (orient == CZT_ROW) {
        if (apply) hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_ROW, true, NLW>), g, b, 0, st, a, p1, p2, tw);
        else hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_ROW, false, NLW>), g, b, 0, st, a, p1, p2, tw);
      } else {
        if (apply) hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_COL, true, NLW>), g, b, 0, st, a, p1, p2, tw);
        else hipLaunchKernelGGL((k_czt_r2r<T, L, CZT_COL, false, NLW>), g, b, 0, st, a, p1, p2, tw);
      }
      return check_launch("czt_r2r");
    }
    static int run(CztR2RArgs a, int orient, bool apply, CztPlan<T> p1, CztPlan<T> p2, int B, hipStream_t st) {
      constexpr int NL0 = CztCfg<T, L>::NL;
      if constexpr (NL0 >= 4) {
        if ((long long)((a.nlines + NL0 - 1) / NL0) * B < 256)
          return go<NL0 / 4>(a, orient, apply, p1, p2, B, st);
      }
      return go<0>(a, orient, apply, p1, p2, B, st);
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

  // ---------------------------------------------------------------- ETD
  // Workspace planes of the any-length ETD chains: the x-half-spectra of the
  // state (Ws), of a stage value (Wg) and of N (Wn), then the separated
  // spectra Uh, Nn, Fh, each B * ny * (nx + 2) complex values.
  struct Planes {
    cpx<T>*Ws, *Wg, *Wn, *Uh, *Nn, *Fh;
  };
  static Planes planes(cpx<T>* work, int B, int ny, int nx) {
    const long long plane = (long long)B * ny * (nx + 2);
    Planes p;
    p.Ws = work;
    p.Wg = p.Ws + plane;
    p.Wn = p.Wg + plane;
    p.Uh = p.Wn + plane;
    p.Nn = p.Uh + plane;
    p.Fh = p.Nn + plane;
    return p;
  }
  struct Ctx {
    CztTables hx, hy;
    int B, ny, nx, nh, sw;
    const InstParams* ip;
    const T* coef;
    hipStream_t st;
    int row_nl(const cpx<T>* in, cpx<T>* out) const {
      return czt_with_len<RowNl>(hx.L, B, ny, nh, in, out, ip, czt_dev<T>(hx), st);
    }
    template <int STAGE>
    int col(const Planes& p, const cpx<T>* in, const cpx<T>* in2, cpx<T>* out) const {
      return czt_with_len<ColEtd<STAGE>::template F>(hy.L, B, nh, sw, in, in2, p.Uh, p.Nn, p.Fh, coef, out,
                                                      czt_dev<T>(hy), st);
    }
  };
  static int ctx(Ctx& c, int B, int ny, int nx, const T* coef, const InstParams* ip, hipStream_t st) {
    int rc = czt_tables(nx, nx, 0, 0, c.hx);
    if (!rc) rc = czt_tables(ny, ny, 0, 0, c.hy);
    c.B = B;
    c.ny = ny;
    c.nx = nx;
    c.nh = nx / 2 + 1;
    c.sw = nx + 2;
    c.ip = ip;
    c.coef = coef;
    c.st = st;
    return rc;
  }
  // ETDRK4 (Cox-Matthews, the four-stage chain of etd_kernels.h): the pack
  // (scaled by 1/nx like row_nl), n_steps steps on the packed state, and
  // the unpack.
  static int etdrk4_pack(int B, int ny, int nx, const T* u, const T* v, long long pitch, long long istride,
                         cpx<T>* work, hipStream_t st) {
    return row_fwd_pack(nx, B, ny, u, v, pitch, istride, planes(work, B, ny, nx).Ws, st, 1.0 / nx);
  }
  static int etdrk4_unpack(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                           cpx<T>* work, hipStream_t st) {
    return row_inv_unpack(nx, B, ny, planes(work, B, ny, nx).Ws, u, v, pitch, istride, ghosts, st, 0.0, nullptr);
  }
  static int etdrk4_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                        const T* coef, const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st) {
    int rc = etdrk4_pack(B, ny, nx, u, v, pitch, istride, work, st);
    if (!rc) rc = etdrk4_advance(B, ny, nx, coef, ip, n_steps, work, st);
    if (rc) return rc;
    return etdrk4_unpack(B, ny, nx, u, v, pitch, istride, ghosts, work, st);
  }
  static int etdrk4_advance(int B, int ny, int nx, const T* coef, const InstParams* ip, int n_steps, cpx<T>* work,
                            hipStream_t st) {
    Ctx c;
    int rc = ctx(c, B, ny, nx, coef, ip, st);
    const Planes p = planes(work, B, ny, nx);
    for (int s = 0; s < n_steps && !rc; ++s) {
      rc = c.row_nl(p.Ws, p.Wn);
      if (!rc) rc = c.template col<1>(p, p.Ws, p.Wn, p.Wg);
      if (!rc) rc = c.row_nl(p.Wg, p.Wn);
      if (!rc) rc = c.template col<2>(p, p.Wn, nullptr, p.Wg);
      if (!rc) rc = c.row_nl(p.Wg, p.Wn);
      if (!rc) rc = c.template col<3>(p, p.Wn, nullptr, p.Wg);
      if (!rc) rc = c.row_nl(p.Wg, p.Wn);
      if (!rc) rc = c.template col<4>(p, p.Wn, nullptr, p.Ws);
    }
    return rc;
  }
  // Exponential Euler (order 1) / AB2 (order 2); prev as etd_exp_run of the
  // power-of-two path (0 none, 1 physical N_{n-1} in nu/nv, 2 the Nn plane).
  static int etd_exp_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                         const T* coef, const InstParams* ip, int order, int prev, const T* nu, const T* nv,
                         int n_steps, cpx<T>* work, hipStream_t st) {
    Ctx c;
    int rc = ctx(c, B, ny, nx, coef, ip, st);
    const Planes p = planes(work, B, ny, nx);
    if (!rc && order == 2 && prev == 1) {
      rc = row_fwd_pack(nx, B, ny, nu, nv, pitch, istride, p.Wg, st, 1.0 / nx);
      if (!rc) rc = c.template col<STAGE_STORE_N>(p, p.Wg, nullptr, nullptr);
    }
    if (!rc) rc = row_fwd_pack(nx, B, ny, u, v, pitch, istride, p.Ws, st, 1.0 / nx);
    cpx<T>* cur = p.Ws;
    for (int s = 0; s < n_steps && !rc; ++s) {
      cpx<T>* next = cur == p.Ws ? p.Wg : p.Ws;
      rc = c.row_nl(cur, p.Wn);
      if (rc) break;
      if (order == 1) rc = c.template col<STAGE_ETD1>(p, cur, p.Wn, next);
      else if (s == 0 && prev == 0) rc = c.template col<STAGE_ETD2_FIRST>(p, cur, p.Wn, next);
      else rc = c.template col<STAGE_ETD2>(p, cur, p.Wn, next);
      cur = next;
    }
    if (rc) return rc;
    return row_inv_unpack(nx, B, ny, cur, u, v, pitch, istride, ghosts, st, 0.0, nullptr);
  }
};

}  // namespace mjx
