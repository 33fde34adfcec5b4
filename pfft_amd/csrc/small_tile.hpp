// small_tile.hpp -- register-resident c2c pass for short mixed-radix
// lengths (NF = 2^a 3^b 5^c <= 128), shared by the per-pass launch
// (kernels_small.hip) and the persistent batch kernel (kernels_batch.hip).
//
// Work split.  For NF <= 16 a workgroup of 256 threads transforms 256
// lines, one line per lane: the lines of the pass are numbered flat over
// the tile dimension and the batch dimensions, and lane `tid` of tile `b`
// takes line 256*b + tid.  The whole line lives in registers, so the FFT
// stages need no LDS and no barriers; every index and every twiddle index
// is a compile-time constant after unrolling (the twiddles are wave-uniform
// loads from the plan's exp(sign*2*pi*i*m/NF) table).  Longer lines are
// split over G lanes (small_split_tile), so that a lane holds at most 16
// elements and the kernel keeps several waves per SIMD.
//
// Coalescing.  When the lines are adjacent along the tile dimension
// (tile stride 1) lanes read element j of 64 consecutive lines together:
// one contiguous run per wave and element.  When the line itself is
// contiguous (element stride 1) the wave stages its 64 lines through LDS in
// 128-byte column chunks: for each chunk the 64 lanes load 64 x 128 bytes
// row segment by row segment (16 lanes per row for complex64), write them
// to a [line][column] buffer padded by one slot, and each lane reads its
// own row back.  Stores mirror the loads.
//
// Stages.  The Stockham stage of radix R over sub-length L maps
//   y[(t/L)*L*R + q*L + k] = sum_r x[t + r*NF/R] W_LR^(r*k) W_R^(r*q),
//   t in [0, NF/R), k = t mod L,
// the recurrence of the general kernel, with the radices taken 8s first
// (4*4 for 16), then 4, 2, 3, 5.
#pragma once
#include "pfk_internal.hpp"

namespace pfk {
namespace small {

template <typename T>
using C2 = cpx<T>;

// radix of stage s for length NF: 8s first, then 4, 2, 3, 5
__host__ __device__ constexpr int radix_at(int nf, int s) {
  int r = 0;
  for (int i = 0; i <= s; ++i) {
    if (nf % 8 == 0 && nf != 16) r = 8;       // 16 = 4*4, not 8*2
    else if (nf % 4 == 0) r = 4;
    else if (nf % 2 == 0) r = 2;
    else if (nf % 3 == 0) r = 3;
    else if (nf % 5 == 0) r = 5;
    else return 0;
    nf /= r;
  }
  return r;
}
__host__ __device__ constexpr int nstages(int nf) {
  int s = 0;
  while (nf > 1) {
    const int r = radix_at(nf, 0);
    if (r == 0) return -1;
    nf /= r;
    ++s;
  }
  return s;
}

template <typename T>
__device__ __forceinline__ C2<T> add(C2<T> a, C2<T> b) { return C2<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> sub(C2<T> a, C2<T> b) { return C2<T>{a.x - b.x, a.y - b.y}; }
// multiply by sg*i (sg = +-1, the transform sign)
template <typename T>
__device__ __forceinline__ C2<T> mul_si(C2<T> a, T sg) { return C2<T>{-sg * a.y, sg * a.x}; }

// in-place DFT-R, R in {2,3,4,5,8}; W_R = exp(sg*2*pi*i/R)
template <typename T, int R>
__device__ __forceinline__ void dft(C2<T> *a, T sg) {
  if constexpr (R == 2) {
    const C2<T> s = add(a[0], a[1]), d = sub(a[0], a[1]);
    a[0] = s;
    a[1] = d;
  } else if constexpr (R == 3) {
    // W_3 = -1/2 + sg*i*sqrt(3)/2
    const T h = (T)0.86602540378443864676372317075293618L;
    const C2<T> b = add(a[1], a[2]), d = sub(a[1], a[2]);
    const C2<T> t{a[0].x - (T)0.5 * b.x, a[0].y - (T)0.5 * b.y};
    const C2<T> u = mul_si(C2<T>{h * d.x, h * d.y}, sg);
    a[0] = add(a[0], b);
    a[1] = add(t, u);
    a[2] = sub(t, u);
  } else if constexpr (R == 4) {
    const C2<T> c0 = add(a[0], a[2]), c1 = sub(a[0], a[2]);
    const C2<T> c2 = add(a[1], a[3]), c3 = mul_si(sub(a[1], a[3]), sg);
    a[0] = add(c0, c2);
    a[2] = sub(c0, c2);
    a[1] = add(c1, c3);
    a[3] = sub(c1, c3);
  } else if constexpr (R == 5) {
    // cos(2pi/5) = (sqrt5-1)/4, cos(4pi/5) = -(sqrt5+1)/4,
    // sin(2pi/5) = sqrt((5+sqrt5)/8), sin(4pi/5) = sqrt((5-sqrt5)/8)
    const T c1 = (T)0.30901699437494742410229341718281906L;
    const T c2 = (T)-0.80901699437494742410229341718281906L;
    const T s1 = (T)0.95105651629515357211643933337938214L;
    const T s2 = (T)0.58778525229247312916870595463907277L;
    const C2<T> b1 = add(a[1], a[4]), b2 = add(a[2], a[3]);
    const C2<T> d1 = sub(a[1], a[4]), d2 = sub(a[2], a[3]);
    const C2<T> t1{a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y};
    const C2<T> t2{a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y};
    const C2<T> u1 = mul_si(C2<T>{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y}, sg);
    const C2<T> u2 = mul_si(C2<T>{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y}, sg);
    a[0] = C2<T>{a[0].x + b1.x + b2.x, a[0].y + b1.y + b2.y};
    a[1] = add(t1, u1);
    a[4] = sub(t1, u1);
    a[2] = add(t2, u2);
    a[3] = sub(t2, u2);
  } else if constexpr (R == 8) {
    C2<T> e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
    dft<T, 4>(e, sg);
    dft<T, 4>(o, sg);
    const T s = (T)0.70710678118654752440084436210484903928L;
    // W8^1 = s(1 + sg i), W8^2 = sg i, W8^3 = s(-1 + sg i)
    const C2<T> t1{s * (o[1].x - sg * o[1].y), s * (o[1].y + sg * o[1].x)};
    const C2<T> t2 = mul_si(o[2], sg);
    const C2<T> t3{s * (-o[3].x - sg * o[3].y), s * (-o[3].y + sg * o[3].x)};
    a[0] = add(e[0], o[0]);
    a[4] = sub(e[0], o[0]);
    a[1] = add(e[1], t1);
    a[5] = sub(e[1], t1);
    a[2] = add(e[2], t2);
    a[6] = sub(e[2], t2);
    a[3] = add(e[3], t3);
    a[7] = sub(e[3], t3);
  }
}

// Stages S.. of the register FFT; L = product of the radices before S.
// tw[TS*m] = exp(sg*2*pi*i*m/NF) (a table of a TS times longer length).
template <typename T, int NF, int S, int L, int TS = 1>
__device__ __forceinline__ void fft(C2<T> *x, const C2<T> *__restrict__ tw, T sg) {
  if constexpr (S < nstages(NF)) {
    constexpr int R = radix_at(NF, S);
    constexpr int NR = NF / R, LR = L * R, M = NF / LR;
    C2<T> y[NF];
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      const int k = t % L;
      C2<T> a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r] = x[t + r * NR];
        if (k != 0 && r != 0) a[r] = cmul(a[r], tw[r * k * M * TS]);
      }
      dft<T, R>(a, sg);
#pragma unroll
      for (int q = 0; q < R; ++q) y[(t / L) * LR + q * L + k] = a[q];
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) x[i] = y[i];
    fft<T, NF, S + 1, LR, TS>(x, tw, sg);
  }
}

// +-1 modulation factor of index j, with selects only (a branch between a
// load and its use makes the compiler wait for each load in turn)
template <typename T>
__device__ __forceinline__ T modsign(const pfk_mod &m, int64_t j) {
  const int64_t t = j + m.start;
  const int s = (m.on && t < m.half) ? ((t & 1) ? -m.extra : m.extra) : 1;
  return (T)s;
}

// complex slots per 128-byte chunk
template <typename T>
constexpr int chunk() { return 128 / (int)sizeof(C2<T>); }

// Shape of the split form (G lanes per line, see small_split_tile).
struct split_shape {
  int g;     // lanes per line (1: one line per lane)
  int m;     // elements per lane = NF / g
  int ni;    // k1 groups per lane in the second step = ceil(m / g)
  int lw;    // lines per wave = 64 / g
  int nw;    // waves per workgroup (4; fewer where 4 waves' LDS exceeds 64 KiB)
  int lpt;   // lines per workgroup tile = nw * lw
  int ib;    // k1 groups per LDS chunk (divides ni)
  int wc;    // chunk width in complex slots (ib * g * g, or NF in one chunk)
  int nch;   // chunks per line (nch * wc = NF)
  int rp;    // LDS row pitch = wc + 1
};

// G for NF <= 128: the smallest of 2, 4, 8 dividing NF with NF/G <= 16
// (failing that the largest dividing one); 1 for NF <= 16.  For NF > 128:
// the power of two up to 64 dividing NF that minimises the larger of the
// two per-lane counts, M and ceil(M/G)*G.  When G divides M the line is
// staged and exchanged in chunks of at most 8 KiB per wave (16 complex64
// or 8 complex128 slots per lane); otherwise in one chunk of NF slots.
__host__ __device__ constexpr split_shape split_of(int nf, bool f64, bool one = false) {
  int g = 1;
  if (nf > 128) {
    int best = 1 << 30;
    for (int c = 2; c <= 64; c *= 2) {
      if (nf % c != 0) continue;
      const int m = nf / c, held = ((m + c - 1) / c) * c;
      const int cost = m > held ? m : held;
      if (cost < best) { best = cost; g = c; }
    }
  } else if (nf > 16) {
    for (int c = 8; c >= 2; c /= 2)
      if (nf % c == 0 && g == 1) g = c;   // largest divisor as the fallback
    for (int c = 2; c <= 8; c *= 2)
      if (nf % c == 0 && nf / c <= 16) { g = c; break; }
  }
  split_shape s{};
  s.g = g;
  s.m = nf / g;
  s.ni = (s.m + g - 1) / g;
  s.lw = 64 / g;
  if (s.m % g == 0 && !one) {
    const int wmax = (f64 ? 8 : 16) * g;
    s.ib = 1;
    for (int d = 1; d <= s.ni; ++d)
      if (s.ni % d == 0 && d * g * g <= wmax) s.ib = d;
    s.wc = s.ib * g * g;
    s.nch = s.ni / s.ib;
  } else {
    s.ib = s.ni;   // one chunk (for an exact split, ni*g*g = nf)
    s.wc = nf;
    s.nch = 1;
  }
  s.rp = s.wc + 1;
  const int csz = f64 ? 16 : 8;
  s.nw = 4;
  while (s.nw > 1 && s.nw * s.lw * s.rp * csz + 2 * s.nw * s.lw * 8 > 65536) s.nw /= 2;
  s.lpt = s.nw * s.lw;
  return s;
}

// LDS bytes one workgroup needs for length nf
__host__ __device__ constexpr size_t lds_bytes(int nf, bool f64, bool one = false) {
  const split_shape s = split_of(nf, f64, one);
  const size_t csz = f64 ? 16 : 8;
  if (s.g == 1) return 4 * 64 * (size_t)(128 / csz + 1) * csz + 2 * 256 * sizeof(int64_t);
  return (size_t)s.nw * s.lw * s.rp * csz + 2 * (size_t)s.lpt * sizeof(int64_t);
}

}  // namespace small

// lines one workgroup tile of the short-length kernel covers
__host__ __device__ constexpr int small_lines_per_tile(int nf, bool f64, bool one = false) {
  return small::split_of(nf, f64, one).lpt;
}

// threads per workgroup of the short-length kernel
__host__ __device__ constexpr int small_threads(int nf, bool f64, bool one = false) {
  return 64 * small::split_of(nf, f64, one).nw;
}

// One tile (256 lines, one per lane) of a c2c pass of length NF (= P.n).
template <typename T, int NF>
__device__ __forceinline__ void small_line_tile(const pfk_pass &P, const int64_t tile,
                                                const T *__restrict__ src, T *__restrict__ dst,
                                                unsigned char *smem) {
  using namespace small;
  using CT = C2<T>;
  constexpr int K = chunk<T>();
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int64_t nlines = P.text;
  for (int i = 0; i < P.nb; ++i) nlines *= P.bext[i];
  const int64_t line = tile * 256 + tid;
  const bool active = line < nlines;
  int64_t sb = -1, db = -1;
  if (active) {
    const int64_t c = line % P.text;
    int64_t bflat = line / P.text;
    sb = c * P.tsin;
    db = c * P.tsout;
    for (int i = 0; i < P.nb; ++i) {
      const int64_t e = P.bext[i];
      const int64_t d = bflat % e;
      bflat /= e;
      sb += d * P.bsin[i];
      db += d * P.bsout[i];
    }
  }
  const CT *s = reinterpret_cast<const CT *>(src);
  CT *d = reinterpret_cast<CT *>(dst);
  CT *chunkbuf = reinterpret_cast<CT *>(smem) + wave * 64 * (K + 1);
  int64_t *sbase = reinterpret_cast<int64_t *>(smem + 4 * 64 * (K + 1) * sizeof(CT));
  int64_t *dbase = sbase + 256;
  const int nin = (int)P.nin, zin = (int)P.zin, nout = (int)P.nout, zout = (int)P.zout;
  const bool lds_in = P.sin == 1 && (P.tsin != 1 || P.text == 1);
  const bool lds_out = P.sout == 1 && (P.tsout != 1 || P.text == 1);
  sbase[tid] = sb;
  dbase[tid] = db;
  lds_barrier();

  // ---------------- load (embed + input modulation) ----------------
  CT x[NF];
  if (lds_in) {
#pragma unroll
    for (int p0 = 0; p0 < NF; p0 += K) {
      const int KC = (NF - p0) < K ? (NF - p0) : K;   // columns in this chunk
      CT v[K];
#pragma unroll
      for (int i = 0; i < K; ++i) {   // issue every load of the chunk first
        if (i < KC) {
          const int e = lane + 64 * i;
          const int row = e / KC, col = e - row * KC;
          const int64_t rb = sbase[wave * 64 + row];
          const int j = p0 + col - zin;
          const bool ok = rb >= 0 && j >= 0 && j < nin;
          v[i] = s[ok ? rb + j : 0];
        }
      }
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (i < KC) {
          const int e = lane + 64 * i;
          const int row = e / KC, col = e - row * KC;
          const int64_t rb = sbase[wave * 64 + row];
          const int j = p0 + col - zin;
          const bool ok = rb >= 0 && j >= 0 && j < nin;
          const T f = modsign<T>(P.imod, j);
          chunkbuf[row * (K + 1) + col] = ok ? CT{v[i].x * f, v[i].y * f} : CT{(T)0, (T)0};
        }
      }
      lds_barrier();
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (q < KC) x[(p0 + q) < NF ? p0 + q : NF - 1] = chunkbuf[lane * (K + 1) + q];
      lds_barrier();
    }
  } else {
    const int64_t sbv = active ? sb : 0;   // inactive lanes read element 0 and drop it
#pragma unroll
    for (int p = 0; p < NF; ++p) {
      const int j = p - zin;
      const bool ok = j >= 0 && j < nin;
      x[p] = s[sbv + (int64_t)(ok ? j : 0) * P.sin];
    }
#pragma unroll
    for (int p = 0; p < NF; ++p) {
      const int j = p - zin;
      const bool ok = active && j >= 0 && j < nin;
      const T f = modsign<T>(P.imod, j);
      x[p] = ok ? CT{x[p].x * f, x[p].y * f} : CT{(T)0, (T)0};
    }
  }

  // ---------------- FFT in registers ----------------
  const T sg = (T)P.sign;
  fft<T, NF, 0, 1>(x, reinterpret_cast<const CT *>(P.tw), sg);

  // ---------------- store (trunc + output modulation) ----------------
  if (lds_out) {
#pragma unroll
    for (int p0 = 0; p0 < NF; p0 += K) {
      const int KC = (NF - p0) < K ? (NF - p0) : K;
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (q < KC) chunkbuf[lane * (K + 1) + q] = x[(p0 + q) < NF ? p0 + q : NF - 1];
      lds_barrier();
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (i < KC) {
          const int e = lane + 64 * i;
          const int row = e / KC, col = e - row * KC;
          const int64_t rb = dbase[wave * 64 + row];
          const int k = p0 + col - zout;
          if (rb >= 0 && k >= 0 && k < nout) {
            CT v = chunkbuf[row * (K + 1) + col];
            const T f = modsign<T>(P.omod, k);
            d[rb + k] = CT{v.x * f, v.y * f};
          }
        }
      }
      lds_barrier();
    }
  } else {
#pragma unroll
    for (int p = 0; p < NF; ++p) {
      const int k = p - zout;
      if (active && k >= 0 && k < nout) {
        const T f = modsign<T>(P.omod, k);
        d[db + (int64_t)k * P.sout] = CT{x[p].x * f, x[p].y * f};
      }
    }
  }
}

// Steps 1-3 of the split form on the lane's registers: x[m] = line[j + G*m]
// (m < M) in, x[i*G + k2] = X[j + G*i + M*k2] out.  buf holds the tile's
// rows (row r of RP slots); the exchange ends with a barrier.
template <typename T, int NF, bool ONE = false>
__device__ __forceinline__ void split_fft(small::C2<T> *x, small::C2<T> *buf, const int r,
                                          const int j, const T sg,
                                          const small::C2<T> *__restrict__ tw) {
  using namespace small;
  using CT = C2<T>;
  constexpr split_shape SH = split_of(NF, sizeof(T) == 8, ONE);
  constexpr int G = SH.g, M = SH.m, NI = SH.ni, IB = SH.ib, NCH = SH.nch, RP = SH.rp;
  constexpr bool EXACT = M % G == 0;
  // ---------------- step 1: M-point DFT of the lane's sequence ----------------
  fft<T, M, 0, 1, G>(x, tw, sg);

  // ---------------- exchange: x[i*G + jj] = Y_jj[j + G*i] ----------------
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
#pragma unroll
    for (int kk = 0; kk < IB * G; ++kk)
      if (c * IB * G + kk < M) buf[r * RP + kk * G + j] = x[c * IB * G + kk];
    lds_barrier();
#pragma unroll
    for (int ii = 0; ii < IB; ++ii) {
      // k1 = j + G*(c*IB + ii); unused slots (k1 >= M) read k1 = 0's values
      int k1l = j + G * ii;
      if (!EXACT && (c * IB + ii + 1) * G > M) k1l = (k1l + c * IB * G < M) ? k1l : 0;
#pragma unroll
      for (int jj = 0; jj < G; ++jj) x[(c * IB + ii) * G + jj] = buf[r * RP + k1l * G + jj];
    }
    lds_barrier();
  }

  // ---------------- twiddle, step 2: x[i*G + k2] = X[j + G*i + M*k2] ----------------
  CT u[G];
#pragma unroll
  for (int jj = 1; jj < G; ++jj) u[jj] = tw[(jj * j) % NF];   // W_NF^(jj*j)
#pragma unroll
  for (int i = 0; i < NI; ++i) {
#pragma unroll
    for (int jj = 1; jj < G; ++jj) {
      CT t = cmul(x[i * G + jj], u[jj]);
      if (i != 0) t = cmul(t, tw[G * ((jj * i) % M)]);   // W_M^(jj*i)
      x[i * G + jj] = t;
    }
    if constexpr (G <= 8)
      dft<T, G>(x + i * G, sg);
    else
      fft<T, G, 0, 1, M>(x + i * G, tw, sg);   // W_G^e = W_NF^(M*e)
  }

}

// One tile (NW*64/G lines) of a c2c pass of length NF = G*M, G lanes per
// line.  Lane j (0 <= j < G) of a line holds the decimated sequence
// x[j + G*m], m in [0, M), and:
//   1. takes its M-point DFT in registers, Y_j[k1] (k1 in [0, M));
//   2. exchanges through LDS so that lane j' holds, for k1 = j' + G*i
//      (i in [0, ceil(M/G))), the G values Y_0[k1] .. Y_{G-1}[k1], and
//      multiplies Y_jj[k1] by W_NF^(jj*k1) = W_NF^(jj*j') W_M^(jj*i);
//   3. takes the G-point DFTs over jj: X[k1 + M*k2] = sum_jj W_G^(jj*k2) (..).
// This is the two-factor split NF = G*M of the DFT (n = j + G*m,
// k = k1 + M*k2) with the first factor over lanes.  Slots with k1 >= M
// (when G does not divide M) are computed and dropped.  Threads take the
// tile's LPT lines line index fastest: thread t is lane j = t / LPT of line
// t % LPT, so that with lines adjacent along the tile dimension each load or
// store instruction of a wave covers runs of min(LPT, 64) adjacent lines
// (the G lanes of a line sit in different waves; the exchange is
// workgroup-wide).  Contiguous lines are staged through LDS in chunks of wc
// columns, each wave staging 64/G rows, as the one-line-per-lane form does;
// the LDS rows are also the exchange buffer of step 2, done in the same
// number of chunks.
template <typename T, int NF>
__device__ __forceinline__ void small_split_tile(const pfk_pass &P, const int64_t tile,
                                                 const T *__restrict__ src, T *__restrict__ dst,
                                                 unsigned char *smem) {
  using namespace small;
  using CT = C2<T>;
  constexpr split_shape SH = split_of(NF, sizeof(T) == 8);
  constexpr int G = SH.g, M = SH.m, NI = SH.ni, LW = SH.lw, LPT = SH.lpt, NW = SH.nw, IB = SH.ib,
                WC = SH.wc, NCH = SH.nch, RP = SH.rp;
  constexpr bool EXACT = M % G == 0;
  constexpr int Q = LW * WC / 64;   // staged slots per lane and chunk (IB*G, or M)
  constexpr int QL = EXACT ? IB * G : M;
  static_assert(G > 1 && G * M == NF && NI * G >= M && NCH * WC == NF && Q * 64 == LW * WC &&
                Q == QL, "shape");
  static_assert(EXACT || NCH == 1, "uneven split in one chunk");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = tid % LPT, j = tid / LPT;   // line within the tile, lane within the line
  const int row0 = wave * LW;               // first line the wave stages
  int64_t nlines = P.text;
  for (int i = 0; i < P.nb; ++i) nlines *= P.bext[i];
  CT *buf = reinterpret_cast<CT *>(smem);   // LPT rows of RP slots
  int64_t *sbase = reinterpret_cast<int64_t *>(smem + NW * LW * RP * sizeof(CT));
  int64_t *dbase = sbase + LPT;
  if (tid < LPT) {
    const int64_t line = tile * LPT + tid;
    int64_t sb = -1, db = -1;
    if (line < nlines) {
      const int64_t c = line % P.text;
      int64_t bflat = line / P.text;
      sb = c * P.tsin;
      db = c * P.tsout;
      for (int i = 0; i < P.nb; ++i) {
        const int64_t e = P.bext[i];
        const int64_t d = bflat % e;
        bflat /= e;
        sb += d * P.bsin[i];
        db += d * P.bsout[i];
      }
    }
    sbase[tid] = sb;
    dbase[tid] = db;
  }
  lds_barrier();
  const int64_t sb = sbase[r], db = dbase[r];
  const bool active = sb >= 0;
  const CT *s = reinterpret_cast<const CT *>(src);
  CT *d = reinterpret_cast<CT *>(dst);
  const int nin = (int)P.nin, zin = (int)P.zin, nout = (int)P.nout, zout = (int)P.zout;
  const bool lds_in = P.sin == 1 && (P.tsin != 1 || P.text == 1);
  const bool lds_out = P.sout == 1 && (P.tsout != 1 || P.text == 1);

  // ---------------- load: x[m] = line[j + G*m] (embed + input modulation) ----------------
  CT x[NI * G];
  if (lds_in) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      CT v[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {   // issue every load of the chunk first
        const int e = lane + 64 * q;
        const int row = e / WC, col = e - row * WC;
        const int64_t rb = sbase[row0 + row];
        const int jj = c * WC + col - zin;
        const bool ok = rb >= 0 && jj >= 0 && jj < nin;
        v[q] = s[ok ? rb + jj : 0];
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int e = lane + 64 * q;
        const int row = e / WC, col = e - row * WC;
        const int64_t rb = sbase[row0 + row];
        const int jj = c * WC + col - zin;
        const bool ok = rb >= 0 && jj >= 0 && jj < nin;
        const T f = modsign<T>(P.imod, jj);
        buf[(row0 + row) * RP + col] = ok ? CT{v[q].x * f, v[q].y * f} : CT{(T)0, (T)0};
      }
      lds_barrier();
#pragma unroll
      for (int mm = 0; mm < QL; ++mm) x[c * QL + mm] = buf[r * RP + j + G * mm];
      lds_barrier();
    }
  } else {
    const int64_t sbv = active ? sb : 0;   // inactive lanes read element 0 and drop it
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int jj = j + G * m - zin;
      const bool ok = jj >= 0 && jj < nin;
      x[m] = s[sbv + (int64_t)(ok ? jj : 0) * P.sin];
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int jj = j + G * m - zin;
      const bool ok = active && jj >= 0 && jj < nin;
      const T f = modsign<T>(P.imod, jj);
      x[m] = ok ? CT{x[m].x * f, x[m].y * f} : CT{(T)0, (T)0};
    }
  }

  split_fft<T, NF>(x, buf, r, j, (T)P.sign, reinterpret_cast<const CT *>(P.tw));

  // ---------------- store (trunc + output modulation) ----------------
  if (lds_out) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int k2 = 0; k2 < G; ++k2) {
          const int base = G * i + M * k2;   // column of lane 0; lane j adds j
          const bool mine = EXACT ? (base >= c * WC && base < (c + 1) * WC) : (j + G * i < M);
          if (mine) buf[r * RP + base - c * WC + j] = x[i * G + k2];
        }
      lds_barrier();
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int e = lane + 64 * q;
        const int row = e / WC, col = e - row * WC;
        const int64_t rb = dbase[row0 + row];
        const int kk = c * WC + col - zout;
        if (rb >= 0 && kk >= 0 && kk < nout) {
          const CT v = buf[(row0 + row) * RP + col];
          const T f = modsign<T>(P.omod, kk);
          d[rb + kk] = CT{v.x * f, v.y * f};
        }
      }
      lds_barrier();
    }
  } else {
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
      for (int k2 = 0; k2 < G; ++k2) {
        const int kk = j + G * i + M * k2 - zout;
        if (active && (EXACT || j + G * i < M) && kk >= 0 && kk < nout) {
          const T f = modsign<T>(P.omod, kk);
          d[db + (int64_t)kk * P.sout] = CT{x[i * G + k2].x * f, x[i * G + k2].y * f};
        }
      }
  }
}

// Line bases of the tile's LPT lines into sbase / dbase (-1 past the end),
// by the first LPT threads; the caller synchronises.
template <int LPT>
__device__ __forceinline__ void tile_bases(const pfk_pass &P, const int64_t tile,
                                           int64_t *sbase, int64_t *dbase) {
  const int tid = threadIdx.x;
  int64_t nlines = P.text;
  for (int i = 0; i < P.nb; ++i) nlines *= P.bext[i];
  if (tid < LPT) {
    const int64_t line = tile * LPT + tid;
    int64_t sb = -1, db = -1;
    if (line < nlines) {
      const int64_t c = line % P.text;
      int64_t bflat = line / P.text;
      sb = c * P.tsin;
      db = c * P.tsout;
      for (int i = 0; i < P.nb; ++i) {
        const int64_t e = P.bext[i];
        const int64_t d = bflat % e;
        bflat /= e;
        sb += d * P.bsin[i];
        db += d * P.bsout[i];
      }
    }
    sbase[tid] = sb;
    dbase[tid] = db;
  }
}

// Visits the (row, col) slots of the tile's LPT rows x ncols columns: with
// the line contiguous (unit element stride) each wave walks its own rows
// column-fastest; otherwise the workgroup walks the slots row-fastest
// (adjacent lines together).
template <int LPT, int LW, int NT, typename F>
__device__ __forceinline__ void tile_slots(const bool col_fast, const int ncols, F &&f) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (col_fast) {
    for (int rr = 0; rr < LW; ++rr)
      for (int col = lane; col < ncols; col += 64) f(wave * LW + rr, col);
  } else {
    const int tot = LPT * ncols;
    for (int idx = tid; idx < tot; idx += NT) f(idx % LPT, idx / LPT);
  }
}

// One tile of an even-length r2c or c2r pass (n = 2*NF, P.nf = NF) in the
// split form; the shape must have one chunk (wc = NF, rows of NF+1 slots).
// r2c: the real line is staged into the LDS rows (real slot p = half of
// complex slot p/2), the lanes take z_m = x_2m + i x_2m+1 at m = j + G*m',
// the split FFT gives Z, which goes back to the rows, and the output slots
// combine Z_k and Z_{NF-k}:
//   X_k = (Z_k + conj Z_{NF-k})/2 + W_n^k (Z_k - conj Z_{NF-k})/(2i).
// c2r: the half spectrum is staged, the lanes form
//   Z_k = E_k + i O_k, E_k = X_k + conj X_{NF-k}, O_k = (X_k - conj X_{NF-k}) W_n^k
// (Z_0 = (X_0 + X_NF, X_0 - X_NF) from the real parts), the split FFT gives
// z, which goes back to the rows, and real output j is slot j of the rows'
// real view.  Embed, trunc, modulation and the Hermitian weights of the
// adjoint passes follow the general kernel.
template <typename T, int NF, int KIND>
__device__ __forceinline__ void small_split_real_tile(const pfk_pass &P, const int64_t tile,
                                                      const T *__restrict__ src,
                                                      T *__restrict__ dst, unsigned char *smem) {
  using namespace small;
  using CT = C2<T>;
  constexpr split_shape SH = split_of(NF, sizeof(T) == 8, true);
  constexpr int G = SH.g, M = SH.m, NI = SH.ni, LW = SH.lw, LPT = SH.lpt, NW = SH.nw,
                RP = SH.rp;
  constexpr int NT = NW * 64;
  static_assert(G > 1 && SH.nch == 1 && SH.wc == NF, "one-chunk split shape");
  const int tid = threadIdx.x;
  const int r = tid % LPT, j = tid / LPT;
  CT *buf = reinterpret_cast<CT *>(smem);
  T *rbuf = reinterpret_cast<T *>(smem);   // real view: row r starts at 2*r*RP
  int64_t *sbase = reinterpret_cast<int64_t *>(smem + NW * LW * RP * sizeof(CT));
  int64_t *dbase = sbase + LPT;
  tile_bases<LPT>(P, tile, sbase, dbase);
  lds_barrier();
  const int nin = (int)P.nin, zin = (int)P.zin, nout = (int)P.nout, zout = (int)P.zout;
  const CT *rtw = reinterpret_cast<const CT *>(P.rtw);

  // ---------------- stage in ----------------
  // Every slot of the tile's rows is written (zero outside the embedded
  // input); all loads of a thread are issued before its LDS writes.  With
  // the line contiguous each wave fills its own rows column-fastest, else
  // the workgroup fills the rows row-fastest.
  const bool col_fast = P.sin == 1;
  const int lane = tid & 63, row0 = (tid >> 6) * LW;
  if constexpr (KIND == PFK_R2C) {
    constexpr int NC = 2 * NF;   // real slots per row
    constexpr int Q = 2 * M;     // = LW*NC/64 = LPT*NC/NT
    static_assert(Q * 64 == LW * NC && Q * NT == LPT * NC, "stage-in split");
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = col_fast ? lane + 64 * q : tid + NT * q;
      const int row = col_fast ? row0 + e / NC : e % LPT;
      const int col = col_fast ? e % NC : e / LPT;
      const int64_t b = sbase[row];
      const int jj = col - zin;
      const bool ok = b >= 0 && jj >= 0 && jj < nin;
      v[q] = src[ok ? b + (int64_t)jj * P.sin : 0];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = col_fast ? lane + 64 * q : tid + NT * q;
      const int row = col_fast ? row0 + e / NC : e % LPT;
      const int col = col_fast ? e % NC : e / LPT;
      const int jj = col - zin;
      const bool ok = sbase[row] >= 0 && jj >= 0 && jj < nin;
      rbuf[2 * row * RP + col] = ok ? v[q] * modsign<T>(P.imod, jj) : (T)0;
    }
  } else {
    const CT *s = reinterpret_cast<const CT *>(src);
    constexpr int NC = NF + 1;                       // complex slots per row
    constexpr int QW = (LW * NC + 63) / 64;          // per lane, wave-own rows
    constexpr int QT = (LPT * NC + NT - 1) / NT;     // per thread, row-fastest
    constexpr int Q = QW > QT ? QW : QT;
    CT v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = col_fast ? lane + 64 * q : tid + NT * q;
      const bool in = col_fast ? (q < QW && e < LW * NC) : (q < QT && e < LPT * NC);
      const int row = in ? (col_fast ? row0 + e / NC : e % LPT) : 0;
      const int col = col_fast ? e % NC : e / LPT;
      const int64_t b = sbase[row];
      const int jj = col - zin;
      const bool ok = in && b >= 0 && jj >= 0 && jj < nin;
      v[q] = s[ok ? b + (int64_t)jj * P.sin : 0];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = col_fast ? lane + 64 * q : tid + NT * q;
      const bool in = col_fast ? (q < QW && e < LW * NC) : (q < QT && e < LPT * NC);
      if (in) {
        const int row = col_fast ? row0 + e / NC : e % LPT;
        const int col = col_fast ? e % NC : e / LPT;
        const int jj = col - zin;
        const bool ok = sbase[row] >= 0 && jj >= 0 && jj < nin;
        T f = modsign<T>(P.imod, jj);
        if (P.hw == 1 && col != 0 && col != NF) f *= (T)0.5;
        buf[row * RP + col] = ok ? CT{v[q].x * f, v[q].y * f} : CT{(T)0, (T)0};
      }
    }
  }
  lds_barrier();

  // ---------------- lane values ----------------
  CT x[NI * G];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int q = j + G * m;
    if constexpr (KIND == PFK_R2C) {
      x[m] = buf[r * RP + q];
    } else {
      const CT a = buf[r * RP + q];
      const CT b = buf[r * RP + NF - q];
      if (q == 0) {
        x[m] = CT{a.x + b.x, a.x - b.x};   // b = X_NF here
      } else {
        const CT E{a.x + b.x, a.y - b.y};
        const CT O = cmul(CT{a.x - b.x, a.y + b.y}, rtw[q]);
        x[m] = CT{E.x - O.y, E.y + O.x};
      }
    }
  }
  lds_barrier();   // the exchange reuses the rows

  split_fft<T, NF, true>(x, buf, r, j, (T)P.sign, reinterpret_cast<const CT *>(P.tw));

#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int k2 = 0; k2 < G; ++k2) {
      const int k1 = j + G * i;
      if (M % G == 0 || k1 < M) buf[r * RP + k1 + M * k2] = x[i * G + k2];
    }
  lds_barrier();

  // ---------------- stage out ----------------
  if constexpr (KIND == PFK_R2C) {
    CT *d = reinterpret_cast<CT *>(dst);
    tile_slots<LPT, LW, NT>(P.sout == 1, nout, [&](int row, int k) {
      const int64_t b = dbase[row];
      if (b < 0) return;
      const int kk = zout + k;
      const CT a = buf[row * RP + (kk == NF ? 0 : kk)];
      const CT c = buf[row * RP + (kk == 0 ? 0 : NF - kk)];
      const CT E{(a.x + c.x) * (T)0.5, (a.y - c.y) * (T)0.5};
      const CT O{(a.y + c.y) * (T)0.5, -(a.x - c.x) * (T)0.5};
      const CT WO = cmul(O, rtw[kk]);
      T f = modsign<T>(P.omod, k);
      if (P.hw == 2 && kk != 0 && kk != NF) f *= (T)2;
      d[b + (int64_t)k * P.sout] = CT{(E.x + WO.x) * f, (E.y + WO.y) * f};
    });
  } else {
    tile_slots<LPT, LW, NT>(P.sout == 1, nout, [&](int row, int k) {
      const int64_t b = dbase[row];
      if (b < 0) return;
      dst[b + (int64_t)k * P.sout] = rbuf[2 * row * RP + zout + k] * modsign<T>(P.omod, k);
    });
  }
}

// Whether the split form computes even-length real passes of length NF
// (nf = NF, n = 2*NF): every split shape, taken in one chunk.
__host__ __device__ constexpr bool split_real_length(int nf, bool f64) {
  return small::split_of(nf, f64, true).g > 1;
}

// One tile of a c2c pass of length NF: small_lines_per_tile(NF) lines.
template <typename T, int NF>
__device__ __forceinline__ void small_tile(const pfk_pass &P, const int64_t tile,
                                           const T *__restrict__ src, T *__restrict__ dst,
                                           unsigned char *smem) {
  if constexpr (small::split_of(NF, sizeof(T) == 8).g > 1)
    small_split_tile<T, NF>(P, tile, src, dst, smem);
  else
    small_line_tile<T, NF>(P, tile, src, dst, smem);
}

// Lengths with a small_tile instantiation: 2^a 3^b 5^c up to 128 except
// the odd ones above 27 (one line per lane would not fit the registers;
// 45 and 90, with 45 values per lane, in complex64 only).
__host__ __device__ constexpr bool small_length(int nf, bool f64) {
  switch (nf) {
    case 2: case 3: case 4: case 5: case 6: case 8: case 9: case 10: case 12: case 15:
    case 16: case 18: case 20: case 24: case 30: case 32: case 36: case 40: case 48:
    case 50: case 54: case 60: case 64: case 72: case 80: case 96: case 100: case 108:
    case 120: case 128:
      return true;
    case 25: case 27:
      return true;
    case 45: case 90:
      return !f64;
    default:
      return false;
  }
}

// Lengths above 128 with a split-form instantiation for the per-pass launch
// (both precisions): 2^a 3^b 5^c, not powers of two, whose split holds at
// most 32 values per lane in either step.
__host__ __device__ constexpr bool mid_length(int nf, bool f64) {
  (void)f64;
  switch (nf) {
    case 144: case 160: case 192: case 240:
    case 200: case 216: case 288: case 320: case 384: case 400: case 432: case 480:
    case 576: case 640: case 768: case 800: case 864: case 960:
      return true;
    default:
      return false;
  }
}

// Runs small_tile for the pass's length (one of small_length's).
template <typename T>
__device__ __forceinline__ void small_dispatch(const pfk_pass &P, const int64_t tile,
                                               const T *src, T *dst, unsigned char *smem) {
  constexpr bool F64 = sizeof(T) == 8;
  switch (P.nf) {
#define PFK_SMALL_CASE(N) \
  case N: if constexpr (small_length(N, F64)) small_tile<T, N>(P, tile, src, dst, smem); break;
    PFK_SMALL_CASE(2) PFK_SMALL_CASE(3) PFK_SMALL_CASE(4) PFK_SMALL_CASE(5) PFK_SMALL_CASE(6)
    PFK_SMALL_CASE(8) PFK_SMALL_CASE(10) PFK_SMALL_CASE(12) PFK_SMALL_CASE(15)
    PFK_SMALL_CASE(16) PFK_SMALL_CASE(20) PFK_SMALL_CASE(24) PFK_SMALL_CASE(30)
    PFK_SMALL_CASE(32) PFK_SMALL_CASE(40) PFK_SMALL_CASE(48) PFK_SMALL_CASE(60)
    PFK_SMALL_CASE(64) PFK_SMALL_CASE(80) PFK_SMALL_CASE(96) PFK_SMALL_CASE(120)
    PFK_SMALL_CASE(128)
    PFK_SMALL_CASE(9) PFK_SMALL_CASE(18) PFK_SMALL_CASE(36) PFK_SMALL_CASE(50)
    PFK_SMALL_CASE(54) PFK_SMALL_CASE(72) PFK_SMALL_CASE(90) PFK_SMALL_CASE(100)
    PFK_SMALL_CASE(108) PFK_SMALL_CASE(25) PFK_SMALL_CASE(27) PFK_SMALL_CASE(45)
#undef PFK_SMALL_CASE
    default: break;
  }
}

}  // namespace pfk
