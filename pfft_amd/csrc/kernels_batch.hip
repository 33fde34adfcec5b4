// kernels_batch.hip -- persistent work-list execution of many passes.
//
// A batch is a list of tasks (one pass of one plan each) grouped by pass
// level: every plan's pass 0, then every plan's pass 1, and so on.  Each
// level runs as one persistent launch over a device work-list: the tiles
// of all tasks of the level form one item sequence, and a fixed grid of
// resident workgroups walks it with a static stride (workgroup b takes
// items b, b + grid, b + 2*grid, ...).  There are no waits and no atomics
// inside a launch; the order between levels (pass p+1 of a plan reads what
// pass p wrote) is given by the launch boundaries.  All level launches are
// captured into one hipGraph and replayed.
#include <stdio.h>
#include <stdlib.h>
#include "pass_tile.hpp"

namespace pfk {

template <typename T>
__global__ void __launch_bounds__(256)
k_batch_level(const pfk_task *__restrict__ tasks, const int t0, const int t1,
              const int64_t item0, const int64_t item1) {
  extern __shared__ unsigned char smem_raw[];
  for (int64_t g = item0 + blockIdx.x; g < item1; g += gridDim.x) {
    int lo = t0, hi = t1 - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].first_item <= g) lo = mid;
      else hi = mid - 1;
    }
    const pfk_task &tk = tasks[lo];
    const int64_t tile = g - tk.first_item;
    const T *src = (const T *)tk.src;
    T *dst = (T *)tk.dst;
    switch (tk.pass.kind) {
      case PFK_C2C: pass_tile<T, PFK_C2C, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      case PFK_R2C: pass_tile<T, PFK_R2C, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      case PFK_C2R: pass_tile<T, PFK_C2R, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
      default: pass_tile<T, PFK_COPY, 16>(tk.pass, tk.C, tk.pitch, tile, src, dst, smem_raw); break;
    }
    __syncthreads();   // LDS is reused by the next item
  }
}

}  // namespace pfk

using namespace pfk;

struct pfk_batch {
  pfk_task *d_tasks;
  int ntasks;
  int nlev;
  int lev_t0[PFK_MAXSTAGES + 1];       /* first task of each level (+ end) */
  int64_t lev_i0[PFK_MAXSTAGES + 1];   /* first item of each level (+ end) */
  size_t lds;
  int grid;
  int prec;
  int use_graph;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

static int batch_launch(pfk_batch *b) {
  for (int l = 0; l < b->nlev; l++) {
    const int64_t items = b->lev_i0[l + 1] - b->lev_i0[l];
    if (items <= 0) continue;
    const int grid = (int)(items < b->grid ? items : b->grid);
    if (b->prec == PFK_F64)
      hipLaunchKernelGGL(k_batch_level<double>, dim3(grid), dim3(256), b->lds, g_stream,
                         b->d_tasks, b->lev_t0[l], b->lev_t0[l + 1], b->lev_i0[l],
                         b->lev_i0[l + 1]);
    else
      hipLaunchKernelGGL(k_batch_level<float>, dim3(grid), dim3(256), b->lds, g_stream,
                         b->d_tasks, b->lev_t0[l], b->lev_t0[l + 1], b->lev_i0[l],
                         b->lev_i0[l + 1]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error("pfk_batch launch", e); return 1; }
  }
  return 0;
}

/* tasks must be grouped by level; task.dep marks the previous pass of the
 * same plan, so a task whose dep lies in the current level starts a new
 * level */
extern "C" pfk_batch *pfk_batch_create(pfk_task *tasks, int ntasks, int use_graph) {
  if (pfk_init() || ntasks <= 0) return nullptr;
  pfk_batch *b = (pfk_batch *)calloc(1, sizeof(pfk_batch));
  b->ntasks = ntasks;
  b->prec = tasks[0].pass.prec;
  b->use_graph = use_graph;
  int64_t item = 0;
  size_t lds = 0;
  b->nlev = 0;
  b->lev_t0[0] = 0;
  b->lev_i0[0] = 0;
  for (int i = 0; i < ntasks; i++) {
    pfk_task *t = &tasks[i];
    if (t->pass.prec != b->prec || t->dep >= i) {
      set_error_msg("pfk_batch: bad task list");
      free(b);
      return nullptr;
    }
    if (t->dep >= b->lev_t0[b->nlev]) {       /* depends on this level: next level */
      if (b->nlev + 1 >= PFK_MAXSTAGES) { set_error_msg("pfk_batch: too many levels"); free(b); return nullptr; }
      b->nlev++;
      b->lev_t0[b->nlev] = i;
      b->lev_i0[b->nlev] = item;
    }
    int C, pitch;
    size_t l;
    if (generic_tiling(&t->pass, &C, &pitch, &l)) { free(b); return nullptr; }
    int64_t nt = (t->pass.text + C - 1) / C;
    for (int k = 0; k < t->pass.nb; k++) nt *= t->pass.bext[k];
    t->C = C;
    t->pitch = pitch;
    t->ntiles = nt;
    t->first_item = item;
    item += nt;
    if (l > lds) lds = l;
  }
  b->nlev++;
  b->lev_t0[b->nlev] = ntasks;
  b->lev_i0[b->nlev] = item;
  b->lds = lds;
  hipError_t e = hipMalloc(&b->d_tasks, sizeof(pfk_task) * ntasks);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_tasks, tasks, sizeof(pfk_task) * ntasks, hipMemcpyHostToDevice);
  if (e != hipSuccess) { set_error("pfk_batch_create", e); pfk_batch_destroy(b); return nullptr; }
  int dev = 0, ncu = 0, occ = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (b->prec == PFK_F64)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_level<double>, 256, lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_level<float>, 256, lds);
  if (e != hipSuccess || occ < 1) occ = 1;
  b->grid = ncu * occ;
  if (use_graph) {
    e = hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const int r = batch_launch(b);
      hipError_t e2 = hipStreamEndCapture(g_stream, &b->graph);
      if (r || e2 != hipSuccess) e = (e2 != hipSuccess) ? e2 : hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { set_error("pfk_batch graph capture", e); pfk_batch_destroy(b); return nullptr; }
  }
  return b;
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
