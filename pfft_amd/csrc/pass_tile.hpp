// pass_tile.hpp -- the general pass computation for one tile of lines,
// shared by the per-pass launch (kernels_generic.hip) and the persistent
// batch kernel (kernels_batch.hip).
#pragma once
#include "pfk_internal.hpp"

namespace pfk {

// +-1 modulation factor of index j, with selects only (a branch between a
// load and its use makes the compiler wait for each load in turn)
__device__ __forceinline__ int mod_sign(const pfk_mod &m, int64_t j) {
  const int64_t t = j + m.start;
  return (m.on && t < m.half) ? ((t & 1) ? -m.extra : m.extra) : 1;
}

// Runs st(i, ld(i)) for i = tid, tid + NT, ... < total, U indices at a
// time with the U loads issued before any is used.  ld(i) must be safe for
// any i in [0, total).
template <int U, typename V, typename Ld, typename St>
__device__ __forceinline__ void batched(const int tid, const int NT, const int total, Ld ld,
                                        St st) {
  for (int b = tid; b < total; b += U * NT) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = b + u * NT;
      v[u] = ld(i < total ? i : b);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = b + u * NT;
      if (i < total) st(i, v[u]);
    }
  }
}

// One tile (C lines) of a pass: the whole pass computation for the lines
// starting at tile `tileid`.  Used by the per-pass launch below and by the
// persistent batch kernel (kernels_batch.hip).  Ends with the stores; the
// caller synchronises before reusing the LDS buffer.
template <typename T, int KIND, int EPT>
__device__ __forceinline__ void pass_tile(const pfk_pass &P, const int C, const int pitch,
                                          const int64_t tileid, const T *__restrict__ src,
                                          T *__restrict__ dst, unsigned char *smem_raw) {
  using C2 = cpx<T>;
  C2 *buf = reinterpret_cast<C2 *>(smem_raw);
  T *rbuf = reinterpret_cast<T *>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;

  const int64_t ntt = (P.text + C - 1) / C;
  const int64_t tt = tileid % ntt;
  int64_t bflat = tileid / ntt;
  const int64_t c0 = tt * C;
  const int nl = (int)((P.text - c0) < C ? (P.text - c0) : C);
  int64_t sb = c0 * P.tsin, db = c0 * P.tsout;
  for (int i = 0; i < P.nb; ++i) {
    const int64_t e = P.bext[i];
    const int64_t d = bflat % e;
    bflat /= e;
    sb += d * P.bsin[i];
    db += d * P.bsout[i];
  }

  const int n = (int)P.n, nf = (int)P.nf, nin = (int)P.nin, zin = (int)P.zin;
  const int nout = (int)P.nout, zout = (int)P.zout;
  const bool ld_tile_fast = (P.tsin == 1 && P.sin != 1);
  const bool st_tile_fast = (P.tsout == 1 && P.sout != 1);

  // ---------------- load (embed + input modulation) ----------------
  if (KIND == PFK_C2C && P.dec_in) {
    // first step of a long line: element j of line cc is index
    // j*dec_in + c0 + cc of the original line (stride tsin), which carries
    // the embed and the input modulation; elements outside the embed read
    // element 0 and are zeroed
    const C2 *s = reinterpret_cast<const C2 *>(src);
    const int64_t line0 = sb - c0 * P.tsin;   // element 0 of the original line
    batched<8, C2>(
        tid, NT, nl * n,
        [&](int idx) {
          const int cc = ld_tile_fast ? idx % nl : idx / n;
          const int j = ld_tile_fast ? idx / nl : idx % n;
          const int64_t jo = (int64_t)j * P.dec_in + c0 + cc - zin;
          const bool ok = jo >= 0 && jo < nin;
          return s[line0 + (ok ? jo : 0) * P.tsin];
        },
        [&](int idx, C2 v) {
          const int cc = ld_tile_fast ? idx % nl : idx / n;
          const int j = ld_tile_fast ? idx / nl : idx % n;
          const int64_t jo = (int64_t)j * P.dec_in + c0 + cc - zin;
          const T f = (jo >= 0 && jo < nin) ? (T)mod_sign(P.imod, jo) : (T)0;
          buf[cc * pitch + j] = C2{v.x * f, v.y * f};
        });
  } else if (KIND == PFK_C2C || KIND == PFK_COPY) {
    const C2 *s = reinterpret_cast<const C2 *>(src);
    if (nin < n)
      for (int idx = tid; idx < nl * n; idx += NT) {
        const int cc = idx / n, p = idx - cc * n;
        if (p < zin || p >= zin + nin) buf[cc * pitch + p] = C2{0, 0};
      }
    batched<8, C2>(
        tid, NT, nl * nin,
        [&](int idx) {
          const int cc = ld_tile_fast ? idx % nl : idx / nin;
          const int j = ld_tile_fast ? idx / nl : idx % nin;
          return s[sb + cc * P.tsin + (int64_t)j * P.sin];
        },
        [&](int idx, C2 v) {
          const int cc = ld_tile_fast ? idx % nl : idx / nin;
          const int j = ld_tile_fast ? idx / nl : idx % nin;
          const T f = (T)mod_sign(P.imod, j);
          buf[cc * pitch + zin + j] = C2{v.x * f, v.y * f};
        });
  } else if (KIND == PFK_R2R) {
    // real line x (length n, embedded at zin) -> extended real sequence u of
    // length nf in complex slots (layout in pfk.h)
    const int kd = P.r2r;
    for (int idx = tid; idx < nl * nf; idx += NT) {
      const int cc = idx / nf, p = idx - cc * nf;
      buf[cc * pitch + p] = C2{0, 0};
    }
    lds_barrier();
    batched<8, T>(
        tid, NT, nl * nin,
        [&](int idx) {
          const int cc = ld_tile_fast ? idx % nl : idx / nin;
          const int j = ld_tile_fast ? idx / nl : idx % nin;
          return src[sb + cc * P.tsin + (int64_t)j * P.sin];
        },
        [&](int idx, T v) {
          const int cc = ld_tile_fast ? idx % nl : idx / nin;
          const int j = ld_tile_fast ? idx / nl : idx % nin;
          const T x = v * (T)mod_sign(P.imod, j);
          const int p = zin + j;
          C2 *L = buf + cc * pitch;
          int a = -1, b = -1;
          T sb_ = (T)1;   // sign of the mirrored entry
          switch (kd) {
            case PFK_DHT: a = p; break;
            case PFK_REDFT00: a = p; if (p > 0 && p < n - 1) b = nf - p; break;
            case PFK_REDFT10: a = 2 * p + 1; b = nf - 2 * p - 1; break;
            case PFK_REDFT01: a = p; if (p > 0) b = nf - p; break;
            case PFK_REDFT11: a = 2 * p + 1; b = nf - 2 * p - 1; break;
            case PFK_RODFT00: a = p + 1; b = nf - p - 1; sb_ = (T)-1; break;
            case PFK_RODFT10: a = 2 * p + 1; b = nf - 2 * p - 1; sb_ = (T)-1; break;
            case PFK_RODFT01: a = p + 1; if (p < n - 1) b = nf - p - 1; sb_ = (T)-1; break;
            default: a = 2 * p + 1; b = nf - 2 * p - 1; sb_ = (T)-1; break;   // RODFT11
          }
          if (a >= 0) L[a] = C2{x, (T)0};
          if (b >= 0) L[b] = C2{sb_ * x, (T)0};
        });
  } else if (KIND == PFK_R2C) {
    if ((n & 1) == 0) {
      // real slot p of line cc is rbuf[2*cc*pitch + p]: complex slot p/2
      for (int idx = tid; idx < nl * n; idx += NT) {
        const int cc = idx / n, p = idx - cc * n;
        if (p < zin || p >= zin + nin) rbuf[2 * cc * pitch + p] = (T)0;
      }
      batched<8, T>(
          tid, NT, nl * nin,
          [&](int idx) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            return src[sb + cc * P.tsin + (int64_t)j * P.sin];
          },
          [&](int idx, T v) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            rbuf[2 * cc * pitch + zin + j] = v * (T)mod_sign(P.imod, j);
          });
    } else {
      for (int idx = tid; idx < nl * n; idx += NT) {
        const int cc = idx / n, p = idx - cc * n;
        if (p < zin || p >= zin + nin) buf[cc * pitch + p] = C2{0, 0};
      }
      batched<8, T>(
          tid, NT, nl * nin,
          [&](int idx) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            return src[sb + cc * P.tsin + (int64_t)j * P.sin];
          },
          [&](int idx, T v) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            buf[cc * pitch + zin + j] = C2{v * (T)mod_sign(P.imod, j), (T)0};
          });
    }
  } else {  // PFK_C2R
    const C2 *s = reinterpret_cast<const C2 *>(src);
    if ((n & 1) == 0) {
      const int h = n / 2;
      if (nin < h + 1)
        for (int idx = tid; idx < nl * (h + 1); idx += NT) {
          const int cc = idx / (h + 1), p = idx - cc * (h + 1);
          if (p < zin || p >= zin + nin) buf[cc * pitch + p] = C2{0, 0};
        }
      batched<8, C2>(
          tid, NT, nl * nin,
          [&](int idx) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            return s[sb + cc * P.tsin + (int64_t)j * P.sin];
          },
          [&](int idx, C2 v) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            T f = (T)mod_sign(P.imod, j);
            const int p = zin + j;
            if (P.hw == 1 && p != 0 && p != h) f *= (T)0.5;
            buf[cc * pitch + p] = C2{v.x * f, v.y * f};
          });
    } else {
      for (int idx = tid; idx < nl * n; idx += NT) {
        const int cc = idx / n, p = idx - cc * n;
        buf[cc * pitch + p] = C2{0, 0};
      }
      lds_barrier();
      batched<8, C2>(
          tid, NT, nl * nin,
          [&](int idx) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            return s[sb + cc * P.tsin + (int64_t)j * P.sin];
          },
          [&](int idx, C2 v) {
            const int cc = ld_tile_fast ? idx % nl : idx / nin;
            const int j = ld_tile_fast ? idx / nl : idx % nin;
            T f = (T)mod_sign(P.imod, j);
            const int p = zin + j;
            if (P.hw == 1 && p != 0) f *= (T)0.5;
            if (p == 0) {
              buf[cc * pitch] = C2{v.x * f, (T)0};
            } else {
              buf[cc * pitch + p] = C2{v.x * f, v.y * f};
              buf[cc * pitch + n - p] = C2{v.x * f, -v.y * f};
            }
          });
    }
  }
  lds_barrier();

  const C2 *tw = reinterpret_cast<const C2 *>(P.tw);
  const C2 *rtw = reinterpret_cast<const C2 *>(P.rtw);

  // ------------- c2r (even n): pack the half spectrum into Z ----------
  if (KIND == PFK_C2R && (n & 1) == 0) {
    const int h = n / 2, hp = h / 2 + 1;
    for (int idx = tid; idx < nl * hp; idx += NT) {
      const int cc = idx / hp, k = idx - cc * hp;
      C2 *L = buf + cc * pitch;
      if (k == 0) {
        const T x0 = L[0].x, xh = L[h].x;
        L[0] = C2{x0 + xh, x0 - xh};
      } else {
        const int k2 = h - k;
        const C2 a = L[k], b = L[k2];
        // E_k = X_k + conj(X_{h-k}),  O_k = (X_k - conj(X_{h-k})) W^k,  Z_k = E_k + i O_k
        {
          const C2 E{a.x + b.x, a.y - b.y};
          const C2 O = cmul(C2{a.x - b.x, a.y + b.y}, rtw[k]);
          L[k] = C2{E.x - O.y, E.y + O.x};
        }
        if (k2 != k) {
          const C2 E{b.x + a.x, b.y - a.y};
          const C2 O = cmul(C2{b.x - a.x, b.y + a.y}, rtw[k2]);
          L[k2] = C2{E.x - O.y, E.y + O.x};
        }
      }
    }
    lds_barrier();
  }

  // ---------------- FFT stages (output-centric Stockham) ----------------
  if (KIND != PFK_COPY) {
    int L = 1;
    for (int s = 0; s < P.nstages; ++s) {
      const int R = P.radix[s];
      const int LR = L * R, M = nf / LR, NR = nf / R;
      C2 res[EPT];
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int og = tid + e * NT;
        if (og < nl * nf) {
          const int cc = og / nf, o = og - cc * nf;
          const int k = o % L, q = (o / L) % R, t = (o / LR) * L + k;
          const C2 *Lp = buf + cc * pitch + t;
          C2 acc = Lp[0];
          const int kq = k + q * L;
          int ex = 0;
          for (int r = 1; r < R; ++r) {
            ex += kq;
            if (ex >= LR) ex -= LR;
            const C2 w = tw[ex * M];
            const C2 v = Lp[r * NR];
            acc.x += v.x * w.x - v.y * w.y;
            acc.y += v.x * w.y + v.y * w.x;
          }
          res[e] = acc;
        }
      }
      lds_barrier();
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int og = tid + e * NT;
        if (og < nl * nf) {
          const int cc = og / nf, o = og - cc * nf;
          buf[cc * pitch + o] = res[e];
        }
      }
      lds_barrier();
      L = LR;
    }
  }

  // ---------------- store (trunc + output modulation) ----------------
  if (KIND == PFK_C2C && (P.dec_out || P.twn)) {
    // steps of a long line.  twn: output k of line cc times
    // W_twn^((c0+cc)*k) = lo[m % 4096] * hi[m / 4096] (m < twn).
    // dec_out: output k of line cc is index k*dec_out + c0 + cc of the
    // original output line (trunc + modulation; sout = dec_out * tsout).
    C2 *d = reinterpret_cast<C2 *>(dst);
    const C2 *lo = reinterpret_cast<const C2 *>(P.tw2);
    const C2 *hi = lo + 4096;
    const int cnt = P.dec_out ? n : nout;
    const int z = P.dec_out ? 0 : zout;
    const int64_t zoff = (int64_t)zout * P.tsout;
    for (int idx = tid; idx < nl * cnt; idx += NT) {
      int cc, k;
      if (st_tile_fast) { cc = idx % nl; k = idx / nl; }
      else { k = idx % cnt; cc = idx / cnt; }
      C2 v = buf[cc * pitch + z + k];
      if (P.twn) {
        const int64_t m = (c0 + cc) * (int64_t)(z + k);
        v = cmul(v, cmul(lo[m & 4095], hi[m >> 12]));
      }
      if (P.dec_out) {
        const int64_t ko = (int64_t)k * P.dec_out + c0 + cc - zout;
        if (ko >= 0 && ko < nout) {
          const T f = (T)mod_sign(P.omod, ko);
          d[db + cc * P.tsout + (int64_t)k * P.sout - zoff] = C2{v.x * f, v.y * f};
        }
      } else {
        const T f = (T)mod_sign(P.omod, k);
        d[db + cc * P.tsout + (int64_t)k * P.sout] = C2{v.x * f, v.y * f};
      }
    }
  } else if (KIND == PFK_C2C || KIND == PFK_COPY) {
    C2 *d = reinterpret_cast<C2 *>(dst);
    for (int idx = tid; idx < nl * nout; idx += NT) {
      int cc, k;
      if (st_tile_fast) { cc = idx % nl; k = idx / nl; }
      else { k = idx % nout; cc = idx / nout; }
      const C2 v = buf[cc * pitch + zout + k];
      const T f = (T)mod_sign(P.omod, k);
      d[db + cc * P.tsout + (int64_t)k * P.sout] = C2{v.x * f, v.y * f};
    }
  } else if (KIND == PFK_R2R) {
    const int kd = P.r2r;
    for (int idx = tid; idx < nl * nout; idx += NT) {
      int cc, k;
      if (st_tile_fast) { cc = idx % nl; k = idx / nl; }
      else { k = idx % nout; cc = idx / nout; }
      const int kk = zout + k;
      const C2 *L = buf + cc * pitch;
      T y;
      switch (kd) {
        case PFK_DHT: y = L[kk].x - L[kk].y; break;
        case PFK_REDFT00: case PFK_REDFT10: y = L[kk].x; break;
        case PFK_REDFT01: case PFK_REDFT11: y = L[2 * kk + 1].x; break;
        case PFK_RODFT00: case PFK_RODFT10: y = -L[kk + 1].y; break;
        default: y = -L[2 * kk + 1].y; break;   // RODFT01, RODFT11
      }
      dst[db + cc * P.tsout + (int64_t)k * P.sout] = y * (T)mod_sign(P.omod, k);
    }
  } else if (KIND == PFK_R2C) {
    C2 *d = reinterpret_cast<C2 *>(dst);
    const int h = n / 2;
    for (int idx = tid; idx < nl * nout; idx += NT) {
      int cc, k;
      if (st_tile_fast) { cc = idx % nl; k = idx / nl; }
      else { k = idx % nout; cc = idx / nout; }
      const int kk = zout + k;
      C2 X;
      if ((n & 1) == 0) {
        const C2 *L = buf + cc * pitch;
        const C2 a = L[kk == h ? 0 : kk];
        const C2 b = L[kk == 0 ? 0 : h - kk];
        // E = (Z_k + conj Z_{h-k})/2 ; O = (Z_k - conj Z_{h-k})/(2i) ; X = E + W^k O
        const C2 E{(a.x + b.x) * (T)0.5, (a.y - b.y) * (T)0.5};
        const C2 O{(a.y + b.y) * (T)0.5, -(a.x - b.x) * (T)0.5};
        const C2 WO = cmul(O, rtw[kk]);
        X = C2{E.x + WO.x, E.y + WO.y};
      } else {
        X = buf[cc * pitch + kk];
      }
      T f = (T)mod_sign(P.omod, k);
      if (P.hw == 2 && kk != 0 && !((n & 1) == 0 && kk == n / 2)) f *= (T)2;
      d[db + cc * P.tsout + (int64_t)k * P.sout] = C2{X.x * f, X.y * f};
    }
  } else {  // C2R
    for (int idx = tid; idx < nl * nout; idx += NT) {
      int cc, k;
      if (st_tile_fast) { cc = idx % nl; k = idx / nl; }
      else { k = idx % nout; cc = idx / nout; }
      const int j = zout + k;
      T x;
      if ((n & 1) == 0) {
        const C2 z = buf[cc * pitch + (j >> 1)];
        x = (j & 1) ? z.y : z.x;
      } else {
        x = buf[cc * pitch + j].x;
      }
      dst[db + cc * P.tsout + (int64_t)k * P.sout] = x * (T)mod_sign(P.omod, k);
    }
  }
}


// Tile width C and LDS row pitch (complex slots) of the general kernel for
// pass p; returns 0 and fills *C, *pitch, *lds_bytes, or nonzero if the
// computed FFT does not fit one workgroup.
int generic_tiling(const pfk_pass *p, int *C, int *pitch, size_t *lds_bytes);

}  // namespace pfk
