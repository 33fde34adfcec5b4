// kernels_batch.hip -- persistent work-list execution of many passes.
//
// A batch is a list of tasks (one pass of one plan each) grouped by pass
// level: every plan's pass 0, then every plan's pass 1, and so on.  Within
// a level the tasks are grouped by the kernel that runs them (the general
// tile kernel, or the short-length kernel of one length), and each group
// runs as one persistent launch over a device work-list: the tiles of the
// group's tasks form one item sequence, and a fixed grid of resident
// workgroups walks it with a static stride (workgroup b takes items b,
// b + grid, b + 2*grid, ...).  There are no waits and no atomics inside a
// launch; the order between levels (pass p+1 of a plan reads what pass p
// wrote) is given by the launch boundaries.  All launches are captured into
// one hipGraph and replayed; within a level the groups are captured as
// parallel branches (forked over four streams, joined before the next
// level), so one group's tail overlaps the next group's start.  Separate
// kernels per group keep each kernel's register allocation (and so its
// occupancy) its own.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pass_tile.hpp"
#include "small_tile.hpp"

namespace pfk {

// task of work-list item g within tasks [t0, t1)
__device__ __forceinline__ int find_task(const pfk_task *__restrict__ tasks, int t0, int t1,
                                         int64_t g) {
  int lo = t0, hi = t1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].first_item <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_batch_generic(const pfk_task *__restrict__ tasks, const int t0, const int t1,
                const int64_t item0, const int64_t item1) {
  extern __shared__ unsigned char smem_raw[];
  for (int64_t g = item0 + blockIdx.x; g < item1; g += gridDim.x) {
    const pfk_task &tk = tasks[find_task(tasks, t0, t1, g)];
    const int64_t tile = g - tk.first_item;
    const T *src = (const T *)tk.src;
    T *dst = (T *)tk.dst;
    switch (tk.pass.kind) {
      case PFK_C2C: pass_tile<T, PFK_C2C, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      case PFK_R2C: pass_tile<T, PFK_R2C, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      case PFK_C2R: pass_tile<T, PFK_C2R, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      case PFK_R2R: pass_tile<T, PFK_R2R, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      default: pass_tile<T, PFK_COPY, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
    }
    lds_barrier();   // LDS is reused by the next item
  }
}

template <typename T, int NF>
__global__ void __launch_bounds__(256)
k_batch_small(const pfk_task *__restrict__ tasks, const int t0, const int t1,
              const int64_t item0, const int64_t item1) {
  static_assert(small_threads(NF, sizeof(T) == 8) == 256, "batch launches run 256 threads");
  extern __shared__ unsigned char smem_raw[];
  for (int64_t g = item0 + blockIdx.x; g < item1; g += gridDim.x) {
    const pfk_task &tk = tasks[find_task(tasks, t0, t1, g)];
    small_tile<T, NF>(tk.pass, g - tk.first_item, (const T *)tk.src, (T *)tk.dst, smem_raw);
    lds_barrier();
  }
}

// short lengths the batch runs with the short-length kernel: those whose
// workgroups have 4 waves (the batch launches 256 threads), odd lengths
// above 16 excepted (one line per lane holds too many registers inside the
// work-list loop)
__host__ __device__ constexpr bool batch_small_length(int nf, bool f64) {
  return small_length(nf, f64) && small_threads(nf, f64) == 256 && !(nf > 16 && (nf & 1));
}

// kernel of a task group: 0 = general, else the short length NF
template <typename T>
static const void *batch_kernel(int cls) {
  if (cls == 0) return reinterpret_cast<const void *>(&k_batch_generic<T>);
  switch (cls) {
#define PFK_BATCH_SMALL(N) \
  case N:                                                                               \
    if constexpr (batch_small_length(N, sizeof(T) == 8))                                \
      return reinterpret_cast<const void *>(&k_batch_small<T, N>);                      \
    break;
    PFK_BATCH_SMALL(2) PFK_BATCH_SMALL(3) PFK_BATCH_SMALL(4) PFK_BATCH_SMALL(5)
    PFK_BATCH_SMALL(6) PFK_BATCH_SMALL(8) PFK_BATCH_SMALL(10) PFK_BATCH_SMALL(12)
    PFK_BATCH_SMALL(15) PFK_BATCH_SMALL(16) PFK_BATCH_SMALL(20) PFK_BATCH_SMALL(24)
    PFK_BATCH_SMALL(30) PFK_BATCH_SMALL(32) PFK_BATCH_SMALL(40) PFK_BATCH_SMALL(48)
    PFK_BATCH_SMALL(60) PFK_BATCH_SMALL(64) PFK_BATCH_SMALL(80) PFK_BATCH_SMALL(96)
    PFK_BATCH_SMALL(120) PFK_BATCH_SMALL(128)
    PFK_BATCH_SMALL(9) PFK_BATCH_SMALL(18) PFK_BATCH_SMALL(36) PFK_BATCH_SMALL(50)
    PFK_BATCH_SMALL(54) PFK_BATCH_SMALL(72) PFK_BATCH_SMALL(90) PFK_BATCH_SMALL(100)
    PFK_BATCH_SMALL(108) PFK_BATCH_SMALL(25) PFK_BATCH_SMALL(27) PFK_BATCH_SMALL(45)
#undef PFK_BATCH_SMALL
    default: break;
  }
  return nullptr;
}

}  // namespace pfk

using namespace pfk;

#define PFK_MAX_GROUPS 256

typedef struct {
  const void *kernel;
  int t0, t1;
  int64_t i0, i1;
  int grid;
  int level;
  size_t lds;
} batch_group;

struct pfk_batch {
  pfk_task *d_tasks;
  int ntasks;
  int ngroups;
  batch_group groups[PFK_MAX_GROUPS];
  int use_graph;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

static int launch_group(const pfk_batch *b, const batch_group *g, hipStream_t s) {
  const int64_t items = g->i1 - g->i0;
  if (items <= 0) return 0;
  const int grid = (int)(items < g->grid ? items : g->grid);
  const pfk_task *tasks = b->d_tasks;
  int t0 = g->t0, t1 = g->t1;
  int64_t i0 = g->i0, i1 = g->i1;
  void *args[] = {(void *)&tasks, (void *)&t0, (void *)&t1, (void *)&i0, (void *)&i1};
  hipError_t e = hipLaunchKernel(g->kernel, dim3(grid), dim3(256), args, g->lds, s);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_batch launch", e); return 1; }
  return 0;
}

static int batch_launch(pfk_batch *b) {
  for (int k = 0; k < b->ngroups; k++)
    if (launch_group(b, &b->groups[k], g_stream)) return 1;
  return 0;
}

// Under graph capture: the groups of one level are independent, so they
// fork over the library stream and three auxiliary streams (default
// priority) and join before the next level; the graph then holds the
// level's launches as parallel branches.
#define PFK_BATCH_AUX 3
static hipStream_t g_aux[PFK_BATCH_AUX];

static int batch_capture_forked(pfk_batch *b) {
  for (int i = 0; i < PFK_BATCH_AUX; i++)
    if (!g_aux[i] && hipStreamCreateWithFlags(&g_aux[i], hipStreamNonBlocking) != hipSuccess) {
      g_aux[i] = nullptr;
      set_error_msg("pfk_batch: auxiliary stream creation failed");
      return 1;
    }
  hipEvent_t ev[PFK_MAX_GROUPS + PFK_MAXSTAGES];
  int nev = 0, err = 0;
  int k = 0;
  while (k < b->ngroups && !err) {
    int kend = k;
    while (kend < b->ngroups && b->groups[kend].level == b->groups[k].level) kend++;
    hipEvent_t start;
    if (hipEventCreateWithFlags(&start, hipEventDisableTiming) != hipSuccess) { err = 1; break; }
    ev[nev++] = start;
    if (hipEventRecord(start, g_stream) != hipSuccess) { err = 1; break; }
    hipEvent_t done[PFK_MAX_GROUPS];
    int ndone = 0;
    for (int i = k; i < kend && !err; i++) {
      const int slot = (i - k) % (PFK_BATCH_AUX + 1);
      hipStream_t s = slot == 0 ? g_stream : g_aux[slot - 1];
      if (s != g_stream && hipStreamWaitEvent(s, start, 0) != hipSuccess) { err = 1; break; }
      if (launch_group(b, &b->groups[i], s)) { err = 1; break; }
      if (s != g_stream && i + PFK_BATCH_AUX + 1 >= kend) {   // last group of this stream
        hipEvent_t d;
        if (hipEventCreateWithFlags(&d, hipEventDisableTiming) != hipSuccess) { err = 1; break; }
        ev[nev++] = d;
        done[ndone++] = d;
        if (hipEventRecord(d, s) != hipSuccess) { err = 1; break; }
      }
    }
    for (int i = 0; i < ndone && !err; i++)
      if (hipStreamWaitEvent(g_stream, done[i], 0) != hipSuccess) err = 1;
    k = kend;
  }
  for (int i = 0; i < nev; i++) (void)hipEventDestroy(ev[i]);
  if (err) set_error_msg("pfk_batch: forked capture failed");
  return err;
}

static int task_class(const pfk_task *t) {
  if (!g_generic_only && t->pass.variant != 1 && t->pass.kind == PFK_C2C &&
      small_covers(&t->pass) && batch_small_length((int)t->pass.nf, t->pass.prec == PFK_F64))
    return (int)t->pass.nf;
  return 0;
}

// Stable counting sort of one level's tasks by kernel class (classes are
// 0 or a short length, below 256), through `tmp` (room for n tasks).
static int sort_level(pfk_task *a, int n, pfk_task *tmp) {
  int start[257];
  memset(start, 0, sizeof(start));
  for (int i = 0; i < n; i++) {
    const int c = task_class(&a[i]);
    if (c < 0 || c > 255) { set_error_msg("pfk_batch: task class out of range"); return 1; }
    start[c + 1]++;
  }
  for (int c = 0; c < 256; c++) start[c + 1] += start[c];
  for (int i = 0; i < n; i++) tmp[start[task_class(&a[i])]++] = a[i];
  memcpy(a, tmp, sizeof(pfk_task) * (size_t)n);
  return 0;
}

/* Groups of a batch (no allocation): tasks must be grouped by level;
 * task.dep marks the previous pass of the same plan, so a task whose dep
 * lies in the current level starts a new level.  The array is reordered
 * (by kernel within each level, through `tmp`, room for ntasks) and its
 * tiling fields filled. */
static int batch_plan_levels(pfk_batch *b, pfk_task *tasks, int ntasks, pfk_task *tmp) {
  const int prec = tasks[0].pass.prec;
  int lev_start[PFK_MAXSTAGES + 1], nlev = 0;
  lev_start[0] = 0;
  for (int i = 0; i < ntasks; i++) {
    if (tasks[i].pass.prec != prec || tasks[i].dep >= i) {
      set_error_msg("pfk_batch: bad task list");
      return 1;
    }
    if (tasks[i].dep >= lev_start[nlev]) {
      if (nlev + 1 >= PFK_MAXSTAGES) { set_error_msg("pfk_batch: too many levels"); return 1; }
      lev_start[++nlev] = i;
    }
  }
  lev_start[++nlev] = ntasks;
  b->ntasks = ntasks;
  b->ngroups = 0;
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  for (int l = 0; l < nlev; l++) {
    const int a = lev_start[l], z = lev_start[l + 1];
    if (sort_level(tasks + a, z - a, tmp)) return 1;
    int i = a;
    while (i < z) {
      const int cls = task_class(&tasks[i]);
      int j = i;
      if (b->ngroups >= PFK_MAX_GROUPS) {
        set_error_msg("pfk_batch: too many task groups");
        return 1;
      }
      batch_group *g = &b->groups[b->ngroups++];
      g->level = l;
      g->t0 = i;
      g->i0 = 0;
      int64_t item = 0;
      size_t lds = 0;
      for (; j < z && task_class(&tasks[j]) == cls; j++) {
        pfk_task *t = &tasks[j];
        int C = -1, pitch = 0;
        size_t ld;
        int64_t nt;
        if (cls) {
          int64_t lines = t->pass.text;
          for (int k = 0; k < t->pass.nb; k++) lines *= t->pass.bext[k];
          const int lpt = small_lines_per_tile(cls, prec == PFK_F64);
          nt = (lines + lpt - 1) / lpt;
          ld = small::lds_bytes(cls, prec == PFK_F64);
        } else {
          if (generic_tiling(&t->pass, &C, &pitch, &ld)) return 1;
          nt = (t->pass.text + C - 1) / C;
          for (int k = 0; k < t->pass.nb; k++) nt *= t->pass.bext[k];
        }
        t->C = C;
        t->pitch = pitch;
        t->ntiles = nt;
        t->first_item = item;
        item += nt;
        if (ld > lds) lds = ld;
      }
      g->t1 = j;
      g->i1 = item;
      g->lds = lds;
      g->kernel = (prec == PFK_F64) ? batch_kernel<double>(cls) : batch_kernel<float>(cls);
      if (!g->kernel) { set_error_msg("pfk_batch: no kernel"); return 1; }
      int occ = 0;
      hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, g->kernel, 256, lds);
      if (e != hipSuccess || occ < 1) occ = 1;
      g->grid = ncu * occ;
      i = j;
    }
  }
  return 0;
}

extern "C" pfk_batch *pfk_batch_create(pfk_task *tasks, int ntasks, int use_graph) {
  if (pfk_init() || ntasks <= 0) return nullptr;
  pfk_batch *b = (pfk_batch *)calloc(1, sizeof(pfk_batch));
  if (!b) return nullptr;
  b->use_graph = use_graph;
  pfk_task *tmp = (pfk_task *)malloc(sizeof(pfk_task) * (size_t)ntasks);
  const int bad = !tmp || batch_plan_levels(b, tasks, ntasks, tmp);
  free(tmp);
  if (bad) { free(b); return nullptr; }
  hipError_t e = hipMalloc(&b->d_tasks, sizeof(pfk_task) * ntasks);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_tasks, tasks, sizeof(pfk_task) * ntasks, hipMemcpyHostToDevice);
  if (e != hipSuccess) { set_error("pfk_batch_create", e); pfk_batch_destroy(b); return nullptr; }
  if (use_graph) {
    e = hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const int r = batch_capture_forked(b);
      hipError_t e2 = hipStreamEndCapture(g_stream, &b->graph);
      if (r || e2 != hipSuccess) e = (e2 != hipSuccess) ? e2 : hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { set_error("pfk_batch graph capture", e); pfk_batch_destroy(b); return nullptr; }
  }
  return b;
}

// Batches of plans that cannot allocate: the device work-list lives in
// this file-scope array instead of an allocation (one copy per GPU, kept
// for the process; PFK_NOALLOC_TASKS passes at most).  Host code uses the
// address hipGetSymbolAddress returns for the current device, never the
// symbol itself.
#define PFK_NOALLOC_TASKS 2048
static __device__ __align__(256) unsigned char scratch[sizeof(pfk_task) * PFK_NOALLOC_TASKS];
#define PFK_MAX_DEVICES 64
static void *g_scratch_ptr[PFK_MAX_DEVICES];
static pfk_batch g_noalloc_batch;                 // group table (host memory)
static pfk_task g_noalloc_tmp[PFK_NOALLOC_TASKS];  // sort buffer (host memory)

extern "C" int pfk_batch_run_noalloc(pfk_task *tasks, int ntasks) {
  if (pfk_init()) return 1;
  if (ntasks <= 0) return 0;
  if (ntasks > PFK_NOALLOC_TASKS) {
    set_error_msg("pfk_batch_run_noalloc: more passes than the fixed work-list holds (2048)");
    return 1;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { set_error("pfk_batch_run_noalloc", e); return 1; }
  if (dev < 0 || dev >= PFK_MAX_DEVICES) {
    set_error_msg("pfk_batch_run_noalloc: device index out of range");
    return 1;
  }
  if (!g_scratch_ptr[dev]) {   // once per device, with that device current
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) { set_error("pfk_batch_run_noalloc: scratch address", e); return 1; }
    g_scratch_ptr[dev] = ptr;
  }
  pfk_batch *b = &g_noalloc_batch;
  b->use_graph = 0;
  if (batch_plan_levels(b, tasks, ntasks, g_noalloc_tmp)) return 1;
  b->d_tasks = (pfk_task *)g_scratch_ptr[dev];
  e = hipMemcpyAsync(b->d_tasks, tasks, sizeof(pfk_task) * (size_t)ntasks, hipMemcpyHostToDevice,
                     g_stream);
  if (e != hipSuccess) { set_error("pfk_batch_run_noalloc: work-list upload", e); return 1; }
  if (batch_launch(b)) return 1;
  e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { set_error("pfk_batch_run_noalloc", e); return 1; }
  return 0;
}

extern "C" int pfk_batch_run(pfk_batch *b) {
  if (!b) return 1;
  hipError_t e;
  if (b->use_graph) {
    e = hipGraphLaunch(b->exec, g_stream);
    if (e != hipSuccess) { set_error("pfk_batch_run(graph)", e); return 1; }
  } else if (batch_launch(b)) {
    return 1;
  }
  e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { set_error("pfk_batch_run", e); return 1; }
  return 0;
}

extern "C" void pfk_batch_destroy(pfk_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->d_tasks);
  free(b);
}
