"""Oracle (test infrastructure only): ghost-cell exchange and reduce.

A numpy restatement of the reference's ghost-cell module on a
one-process mesh, step by step in the reference's order.  Only tests/
use it.

Followed (reference file:line):
  evaluate_user_gcells               util/util.c:85-94 (NULL -> no ghost cells)
  local_size_gc_internal             gcell/gcells_plan.c:48-74
  plan_many_rgc / plan_many_cgc      api/api-adv.c:232-326 (padded real size, r2c complex
                                     size, complex data as 2*howmany reals)
  plan_rgc_internal                  gcell/gcells_plan.c:77-200 (transposed order, blocks,
                                     local sizes, NULL for zero ghost cells, rank 3 only)
  vcopy_INT_transposed               gcell/gcells_plan.c:222-235
  exchange_gc / reduce_gc            gcell/gcells_plan.c:373-405
  pad_array_with_zeros_outside_3d    gcell/gcells_plan.c:408-441
  truncate_array_outside_3d          gcell/gcells_plan.c:444-476
  exchange below / above             gcell/gcells_sendrecv.c:149-152, 167-255, 258-330
  reduce above / below               gcell/gcells_sendrecv.c:425-478, 481-536
  local_block_size_shifted           kernel/block.c:42-51: the whole dimension on one process
On one process every neighbour is the process itself: a send/recv of
slices is a copy of packed slices (all read, then all written), an
add-send/recv adds them (gcells_sendrecv.c:543-596).
"""
import numpy as np

GC_TRANSPOSED = 1 << 0
GC_SENDRECV = 1 << 1
GC_RMA = 1 << 2
GC_R2C = 1 << 3
GC_PADDED = 1 << 4
GC_PADDED_R2C = GC_R2C | GC_PADDED


def _gcells(rnk, gc):
    return [0] * rnk if gc is None else [int(v) for v in gc[:rnk]]


def local_size_gc(local_n, local_start, howmany, gc_below, gc_above):
    """gcells_plan.c:48-74: (memory in tuples of `howmany`, local_ngc,
    local_gc_start)."""
    rnk = len(local_n)
    gb, ga = _gcells(rnk, gc_below), _gcells(rnk, gc_above)
    ngc = [gb[t] + local_n[t] + ga[t] for t in range(rnk)]
    start = [local_start[t] - gb[t] for t in range(rnk)]
    mem = howmany
    for t in range(rnk):
        mem *= ngc[t]
    return mem, ngc, start


def _transposed(v, rnk_pm, flags):
    """gcells_plan.c:222-235"""
    cp = list(v)
    for t in range(rnk_pm):
        cp[t] = v[(t + 1) % rnk_pm] if (flags & GC_TRANSPOSED) else v[t]
    return cp


def plan_params(n, howmany, gc_below, gc_above, flags, rnk_pm, complex_data, block=None):
    """Geometry of a ghost-cell plan (api-adv.c:232-326, gcells_plan.c:77-200)
    on a one-process mesh of dimension rnk_pm; None where the reference
    returns a NULL plan.  Sizes are in the plan's (transposed) order, the
    tuple counts reals."""
    rnk = len(n)
    pn = list(n)
    tuple_ = howmany
    if complex_data:
        flags &= ~GC_PADDED
        if flags & GC_R2C:
            pn[rnk - 1] = n[rnk - 1] // 2 + 1
            flags &= ~GC_R2C
        tuple_ = 2 * howmany
    elif flags & GC_PADDED:
        pn[rnk - 1] = 2 * (n[rnk - 1] // 2 + 1)
    gb, ga = _gcells(rnk, gc_below), _gcells(rnk, gc_above)
    if all(v == 0 for v in gb + ga):
        return None
    if rnk != 3:
        return None
    tn = _transposed(pn, rnk_pm, flags)
    tgb = _transposed(gb, rnk_pm, flags)
    tga = _transposed(ga, rnk_pm, flags)
    # evaluate_user_block_size on one process: the default block is the
    # whole dimension; decompose() gives the first block
    loc = []
    for t in range(rnk):
        b = tn[t] if (t >= rnk_pm or block is None or block[t] == 0) else block[t]
        loc.append(min(b, tn[t]))
    ngc = [tgb[t] + loc[t] + tga[t] for t in range(rnk)]
    return {"n": tn, "loc_n": loc, "gb": tgb, "ga": tga, "ngc": ngc, "tuple": tuple_}


def _box(buf, shape, tuple_):
    cnt = tuple_ * int(np.prod(shape))
    return buf[:cnt].reshape(list(shape) + [tuple_])


def _slices(dim, start, count):
    sl = [slice(None)] * 4
    sl[dim] = slice(start, start + count)
    return tuple(sl)


def pad_zeros_outside(buf, loc_n, gb, ga, tuple_):
    """gcells_plan.c:408-441: move the local block into the enlarged box,
    zeros outside it."""
    no = [gb[t] + loc_n[t] + ga[t] for t in range(3)]
    if list(no) == list(loc_n):
        return buf
    out = buf.copy()
    src = _box(buf, loc_n, tuple_).copy()
    box = np.zeros(no + [tuple_], dtype=buf.dtype)
    box[gb[0]:gb[0] + loc_n[0], gb[1]:gb[1] + loc_n[1], gb[2]:gb[2] + loc_n[2], :] = src
    out[:box.size] = box.reshape(-1)
    return out


def truncate_outside(buf, ngc, gb, ga, tuple_):
    """gcells_plan.c:444-476: compact the local block to the front, zeros
    up to the size of the box."""
    no = [ngc[t] - gb[t] - ga[t] for t in range(3)]
    if list(no) == list(ngc):
        return buf
    out = buf.copy()
    box = _box(buf, ngc, tuple_)
    blk = box[gb[0]:gb[0] + no[0], gb[1]:gb[1] + no[1], gb[2]:gb[2] + no[2], :].copy()
    total = tuple_ * int(np.prod(ngc))
    out[:total] = 0
    out[:blk.size] = blk.reshape(-1)
    return out


def _copy_slices(box, dim, send, recv, count):
    if count <= 0:
        return
    packed = box[_slices(dim, send, count)].copy()
    box[_slices(dim, recv, count)] = packed


def _add_slices(box, dim, send, recv, count):
    if count <= 0:
        return
    packed = box[_slices(dim, send, count)].copy()
    box[_slices(dim, recv, count)] += packed


def exchange(buf, p):
    """pfft_exchange on one process: `buf` holds the local block (loc_n x
    tuple reals) at its front; returns the buffer with the ghost-cell box."""
    buf = pad_zeros_outside(buf, p["loc_n"], p["gb"], p["ga"], p["tuple"])
    box = _box(buf, p["ngc"], p["tuple"])
    for dim in range(3):
        N = p["n"][dim]          # local_block_size_shifted at any shift
        start = p["gb"][dim]
        end = start + p["loc_n"][dim]
        # below: gcells_sendrecv.c:258-330
        send, recv, left = end, start, p["gb"][dim]
        while left > 0:
            cur = min(left, N)
            send -= cur
            recv -= cur
            _copy_slices(box, dim, send, recv, cur)
            left -= cur
        # above: gcells_sendrecv.c:167-255
        send, recv, left = start, end, p["ga"][dim]
        while left > 0:
            cur = min(left, N)
            _copy_slices(box, dim, send, recv, cur)
            send += cur
            recv += cur
            left -= cur
    return buf


def reduce(buf, p):
    """pfft_reduce on one process: the adjoint of exchange.  `buf` holds
    the ghost-cell box; returns the buffer with the local block at its
    front and zeros up to the size of the box."""
    buf = buf.copy()
    box = _box(buf, p["ngc"], p["tuple"])
    for dim in range(3):
        N = p["n"][dim]
        start = p["gb"][dim]
        end = start + p["loc_n"][dim]
        # below: gcells_sendrecv.c:481-536
        gb = p["gb"][dim]
        nsh = -((gb + N - 1) // N) if gb > 0 else 0
        possible = -nsh * N
        send, recv, left = 0, end - gb, gb
        for _ in range(nsh + 1, 1):
            possible -= N
            cur = left - min(left, possible)
            _add_slices(box, dim, send, recv, cur)
            send += cur
            recv += cur
            left -= cur
        # above: gcells_sendrecv.c:425-478
        ga = p["ga"][dim]
        nsh = (ga + N - 1) // N if ga > 0 else 0
        possible = nsh * N
        send, recv, left = end + ga, start + ga, ga
        for _ in range(nsh - 1, -1, -1):
            possible -= N
            cur = left - min(left, possible)
            send -= cur
            recv -= cur
            _add_slices(box, dim, send, recv, cur)
            left -= cur
    return truncate_outside(buf, p["ngc"], p["gb"], p["ga"], p["tuple"])
