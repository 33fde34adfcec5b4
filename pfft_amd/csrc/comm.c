/*
 * comm.c -- process meshes on one process.
 *
 * The library runs one process on one GPU, so every communicator it sees
 * has one rank.  What still matters for the layout contract is the
 * Cartesian dimension of the communicator handed to local_size/plan:
 * a 2-d mesh gives the [k1][k2][k0] transposed order, a 1-d mesh (or a
 * plain, non-Cartesian communicator, which the reference turns into a
 * 1-d mesh: kernel/procmesh.c:114-129) gives [k1][k0][k2].
 *
 * pfft_create_procmesh* hand out fresh integer handles recorded in a small
 * registry; any other handle (MPI_COMM_WORLD, MPI_COMM_SELF, a handle from
 * a real MPI library) is a one-rank non-Cartesian communicator.
 *
 * Reference behaviour restated (kernel/procmesh.c:36-89):
 *   - returns 1 if the product of np differs from the communicator size;
 *   - otherwise creates the Cartesian communicator and returns 0
 *     (MPI_SUCCESS).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "host.h"

#define PA_MAXCOMMS 4096
#define PA_COMM_BASE 0x5FA00000

typedef struct {
  int used;
  int rnk;
  int np[PA_MAXRNK];
} cart_entry;

static cart_entry g_carts[PA_MAXCOMMS];
static int g_ncarts = 0;

int pa_mesh_of(MPI_Comm comm, pa_mesh *m) {
  memset(m, 0, sizeof(*m));
  const long idx = (long)comm - PA_COMM_BASE;
  if (idx >= 0 && idx < g_ncarts && g_carts[idx].used) {
    m->rnk_pm = g_carts[idx].rnk;
    m->size = 1;
    for (int t = 0; t < m->rnk_pm; t++) { m->np[t] = g_carts[idx].np[t]; m->size *= m->np[t]; }
    return 0;
  }
  if (comm == MPI_COMM_NULL) return 1;
  /* non-Cartesian communicator with one rank: assure_cart_comm makes it a
   * 1-d mesh of size 1 (kernel/procmesh.c:114-129) */
  m->rnk_pm = 1;
  m->np[0] = 1;
  m->size = 1;
  return 0;
}

int pa_create_procmesh(int rnk, MPI_Comm comm, const int *np, MPI_Comm *comm_cart) {
  pa_mesh parent;
  if (rnk < 1 || rnk > PA_MAXRNK || !np) return 1;
  if (pa_mesh_of(comm, &parent)) return 1;
  long prod = 1;
  for (int t = 0; t < rnk; t++) prod *= np[t];
  if (prod != parent.size) return 1;       /* procmesh.c:47-48 */
  if (g_ncarts >= PA_MAXCOMMS) return 1;
  cart_entry *e = &g_carts[g_ncarts];
  e->used = 1;
  e->rnk = rnk;
  for (int t = 0; t < rnk; t++) e->np[t] = np[t];
  *comm_cart = (MPI_Comm)(PA_COMM_BASE + g_ncarts);
  g_ncarts++;
  return 0;
}
