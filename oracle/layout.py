"""Oracle (test infrastructure only): integer layout contract of PFFT.

A literal, function-by-function restatement in pure Python of the
reference's local-size machinery on a process mesh whose dimensions all
have one process (coordinate 0).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline may use this module; the product never does.

Reference functions followed (file:line in the reference tree):
  local_size_partrafo          kernel/partrafo.c:202-314
  local_block_partrafo         kernel/partrafo.c:99-199
  evaluate_blocks              kernel/partrafo.c:652-701
  init_param_local_size        kernel/partrafo.c:703-732
  init_param_size_and_trafo_flags  kernel/partrafo.c:735-834
  local_size_partrafo_transposed   kernel/partrafo-transposed.c:88-182
  local_block_partrafo_transposed  kernel/partrafo-transposed.c:53-86
  initialize_Nb / calculate_tuple_size  partrafo-transposed.c:294-337
  local_size_transposed / decompose_*   partrafo-transposed.c:341-411
  get_outrafo_param / local_size_outrafo  kernel/outrafo.c:46-115
  local_size_ousam_dd          kernel/ousample.c:55-75
  local_size_sertrafo          kernel/sertrafo.c:152-169
  get_global_transp_param / local_size_global_transp  kernel/transpose.c:63-137
  block arithmetic             kernel/block.c:31-95, util/util.c:43-83,207-227
Third-party: FFTW 3.3.4 (un-vendored; pinned by conf/travis-install-fftw.sh:4)
XM(local_size_many_transposed) on one process, restated from FFTW's
published block semantics in fftw_local_size_many_transposed() below.

Plan validation (which flag/size combinations give a NULL plan) follows
PX(plan_partrafo) kernel/partrafo.c:337-373 in plan_is_null().
"""

# flag values: api/pfft.h:528-545
TRANSPOSED_IN = 1 << 0
TRANSPOSED_OUT = 1 << 1
SHIFTED_IN = 1 << 2
SHIFTED_OUT = 1 << 3
ESTIMATE = 1 << 4
PRESERVE_INPUT = 1 << 8
DESTROY_INPUT = 1 << 9
PADDED_R2C = 1 << 11
PADDED_C2R = PADDED_R2C
DEFAULT_BLOCK = 0

# internal trafo flags: kernel/ipfft.h:137-163
T_C2C = 1 << 0
T_R2C = 1 << 1
T_C2R = 1 << 2
T_R2R = 0x7F8
T_SKIP = 1 << 11
T_PHANTOM = 1 << 12
T_PADDED = 1 << 13
T_RDFT = T_R2C | T_C2R


def physical_dft_size_1d(n, tf):
    return n // 2 + 1 if tf & T_RDFT else n


def physical_dft_size(n, tf):
    pn = list(n)
    pn[-1] = physical_dft_size_1d(n[-1], tf)
    return pn


def prod(v):
    p = 1
    for x in v:
        p *= x
    return p


def cdiv(a, b):
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def cmod(a, b):
    return a - b * cdiv(a, b)


# ---- kernel/block.c ----
def num_blocks(N, blk):
    return (N + blk - 1) // blk


def local_block_offset(N, blk, which):
    return 0 if which >= num_blocks(N, blk) else which * blk


def local_block_size(N, blk, which):
    b = num_blocks(N, blk)
    if which >= b:
        return 0
    return N - which * blk if which == b - 1 else blk


def global_block_size(N, user_blk, P):
    return (N + P - 1) // P if user_blk == DEFAULT_BLOCK else user_blk


def evaluate_user_block_size(r, pn, block, np_pm):
    return [global_block_size(pn[t], DEFAULT_BLOCK if block is None else block[t], np_pm[t])
            for t in range(r)]


def decompose(pn, blk, r, coords, ln, ls, off):
    for t in range(r):
        ls[t + off] = local_block_offset(pn[t + off], blk[t], coords[t])
        ln[t + off] = local_block_size(pn[t + off], blk[t], coords[t])


def decompose_nontransposed(n, blk, r, coords, tf):
    pn = physical_dft_size(n, tf)
    ln, ls = list(pn), [0] * len(n)
    decompose(pn, blk, r, coords, ln, ls, 0)
    return ln, ls


def decompose_transposed(n, blk, r, coords, tf):
    pn = physical_dft_size(n, tf)
    ln, ls = list(pn), [0] * len(n)
    decompose(pn, blk, r, coords, ln, ls, 1)
    return ln, ls


def local_size_transposed(ni, no, iblock, oblock, r, coords, tf, transp):
    if transp & TRANSPOSED_OUT:
        lni, lis = decompose_nontransposed(ni, iblock, r, coords, tf)
        lno, los = decompose_transposed(no, oblock, r, coords, tf)
    else:
        lni, lis = decompose_transposed(ni, iblock, r, coords, tf)
        lno, los = decompose_nontransposed(no, oblock, r, coords, tf)
    return lni, lis, lno, los


def initialize_Nb(r, lni, lno, transp):
    ln = lno if transp & TRANSPOSED_IN else lni
    return prod(ln[:r])


def calculate_tuple_size(rnk_n, ni, no, r, howmany, tf, transp):
    n = no if transp & TRANSPOSED_OUT else ni
    tpl = howmany
    if r + 1 < rnk_n:
        tpl *= physical_dft_size_1d(n[rnk_n - 1], tf)
    for t in range(r + 1, rnk_n - 1):
        tpl *= n[t]
    return tpl


def local_size_ousam_dd(nb, ni, no, hm, tf):
    return max(nb * prod(physical_dft_size(ni, tf)) * hm, nb * prod(physical_dft_size(no, tf)) * hm)


def local_size_sertrafo(nb, n, hm, tf):
    return nb * prod(physical_dft_size(n, tf)) * hm


def local_size_outrafo(nb, n, ni, no, hm, tf):
    mem = 1
    mem = max(mem, local_size_ousam_dd(nb, ni, n, hm, tf))
    mem = max(mem, local_size_sertrafo(nb, n, hm, tf))
    mem = max(mem, local_size_ousam_dd(nb, n, no, hm, tf))
    return mem


def get_outrafo_param(step, r, n, ni, no, lni, lno, transp, tfs):
    if transp & TRANSPOSED_IN:
        lni, lno = lno, lni
    Nb = 1
    for t in range(step):
        Nb *= lno[r - t]
    for t in range(r - step):
        Nb *= lni[t]
    return Nb, n[r - step], ni[r - step], no[r - step], tfs[r - step]


def _gtransp_core(step, r, ni, no, lni, lno, tpl, iblock, oblock, tf):
    N0 = no[r - step]
    N1 = ni[r - 1 - step]
    h0 = 1
    for t in range(step):
        h0 *= lno[r - t]
    for t in range(r - 1 - step):
        h0 *= lni[t]
    hm = tpl * (1 if tf & T_R2R else 2)
    return N0, N1, h0, 1, hm, oblock[r - 1 - step], iblock[r - 1 - step]


def get_global_transp_param(step, r, ni, no, lni, lno, tpl, iblock, oblock, tf, transp):
    if transp & TRANSPOSED_IN:
        a = _gtransp_core(step, r, no, ni, lno, lni, tpl, oblock, iblock, tf)
        # outputs assigned to (N1, N0, h1, h0, hm, blk1, blk0)
        N1, N0, h1, h0, hm, blk1, blk0 = a
    else:
        N0, N1, h0, h1, hm, blk0, blk1 = _gtransp_core(step, r, ni, no, lni, lno, tpl, iblock,
                                                       oblock, tf)
    return N0, N1, h0, h1, hm, blk0, blk1


def fftw_local_size_many_transposed(N, hm, b0, b1, P=1):
    """FFTW-MPI local_size_many_transposed for a 2-d array N[0] x N[1] on
    process 0 of P: input block min(b0, N0) rows, transposed output block
    min(b1, N1) rows; returns hm * max(input, output) element count."""
    l0 = min(b0, N[0]) if b0 > 0 else N[0]
    l1 = min(b1, N[1]) if b1 > 0 else N[1]
    return hm * max(l0 * N[1], l1 * N[0])


def local_size_global_transp(N0, N1, h0, h1, hm, blk0, blk1):
    N = [N1 * h1, N0 * h0]
    blk = [blk1 * h1, blk0 * h0]
    if N[0] * N[1] * h0 * h1 * hm == 0:
        return 1
    return max(fftw_local_size_many_transposed(N, hm, blk[0], blk[1]), 1)


def overwrite_real_counts(ni, no, tf, lni, lno):
    if tf & T_R2C:
        if tf & T_PADDED:
            lni[-1] *= 2
        else:
            lni[-1] = ni[-1]
    if tf & T_C2R:
        if tf & T_PADDED:
            lno[-1] *= 2
        else:
            lno[-1] = no[-1]


def local_size_partrafo_transposed(n, ni, no, howmany, iblock, oblock, r, coords, transp, tfs):
    rnk_n = len(n)
    tf = tfs[r]
    mem = 1
    lni, lis, lno, los = local_size_transposed(ni, no, iblock, oblock, r, coords, tf, transp)
    Nb = initialize_Nb(r, lni, lno, transp)
    tpl = calculate_tuple_size(rnk_n, ni, no, r, howmany, tf, transp)
    mem = max(mem, local_size_outrafo(Nb, n[r:], ni[r:], no[r:], howmany, tf))
    pni = physical_dft_size(ni, tf)
    pno = physical_dft_size(no, tf)
    if tf & T_RDFT:
        tf = T_C2C
    for t in range(r):
        N0, N1, h0, h1, hm, b0, b1 = get_global_transp_param(t, r, pni, pno, lni, lno, tpl, iblock,
                                                            oblock, tf, transp)
        hm = 1
        mem = max(mem, local_size_global_transp(N0, N1, h0, h1, hm, b0, b1))
        Nb, N, Ni, No, tf = get_outrafo_param(t + 1, r, n, ni, no, lni, lno, transp, tfs)
        mem = max(mem, local_size_outrafo(Nb, [N], [Ni], [No], tpl, tf))
    overwrite_real_counts(ni, no, tfs[r], lni, lno)
    return mem, lni, lis, lno, los


def get_skip_flag(skip, t):
    return 0 if skip is None else (T_SKIP if skip[t] else 0)


def init_param_size_and_trafo_flags(n, ni, no, r, tf, transp, skip):
    rnk_n = len(n)
    tfs_to = [tf | get_skip_flag(skip, t) for t in range(r + 1)]
    ni_to, n_to, no_to = list(ni), list(n), list(no)
    tfs_ti = [tf | T_SKIP for t in range(r + 1)]
    ni_ti, n_ti, no_ti = list(no), list(no), list(no)
    if transp & TRANSPOSED_IN:
        tfs_ti = [tf | get_skip_flag(skip, t) for t in range(r + 1)]
        ni_ti, n_ti, no_ti = list(ni), list(n), list(no)
    if tf & T_R2C:
        tfs_to = [(tf if t == r else T_C2C) | get_skip_flag(skip, t) for t in range(r + 1)]
        ni_to, n_to, no_to = list(ni), list(n), list(no)
        tfs_ti = [T_C2C | T_SKIP for t in range(r + 1)]
        for t in range(rnk_n - 1):
            ni_ti[t] = n_ti[t] = no_ti[t] = no[t]
        ni_ti[-1] = n_ti[-1] = no_ti[-1] = no[-1] // 2 + 1
    if tf & T_C2R:
        tfs_to = [T_C2C | T_SKIP for t in range(r + 1)]
        for t in range(rnk_n - 1):
            ni_to[t] = n_to[t] = no_to[t] = ni[t]
        ni_to[-1] = n_to[-1] = no_to[-1] = ni[-1] // 2 + 1
        tfs_ti = [(tf if t == r else T_C2C) | get_skip_flag(skip, t) for t in range(r + 1)]
        ni_ti, n_ti, no_ti = list(ni), list(n), list(no)
    if transp & TRANSPOSED_OUT:
        tfs_ti = [0] * (r + 1)
        ni_ti, n_ti, no_ti = [0] * rnk_n, [0] * rnk_n, [0] * rnk_n
    if transp & TRANSPOSED_IN:
        # partrafo.c:815 assigns to the local pointer: the to-flags survive
        ni_to, n_to, no_to = [0] * rnk_n, [0] * rnk_n, [0] * rnk_n
    # partrafo.c:828: `~transp_flag & (T_IN || T_OUT)` tests bit 0 only
    if ~transp & (1 if (TRANSPOSED_IN or TRANSPOSED_OUT) else 0):
        if tf & T_C2R:
            tfs_to[0] = T_PHANTOM
        else:
            tfs_ti[0] = T_PHANTOM
    return (n_to, ni_to, no_to, tfs_to), (n_ti, ni_ti, no_ti, tfs_ti)


def evaluate_blocks(ni, no, iblock_user, oblock_user, r, np_pm, tf, transp):
    pni = physical_dft_size(ni, tf)
    pno = physical_dft_size(no, tf)
    pnm = pni if (tf & T_C2R) or (transp & TRANSPOSED_IN) else pno
    mblock_user = None
    if transp & TRANSPOSED_IN:
        mblock_user = iblock_user
    if transp & TRANSPOSED_OUT:
        mblock_user = oblock_user
    iblk = [0] * r
    mblk = [0] * r
    oblk = [0] * r
    if not transp & TRANSPOSED_IN:
        iblk = evaluate_user_block_size(r, pni, iblock_user, np_pm)
    mblk = evaluate_user_block_size(r, pnm[1:], mblock_user, np_pm)
    if not transp & TRANSPOSED_OUT:
        oblk = evaluate_user_block_size(r, pno, oblock_user, np_pm)
    return iblk, mblk, oblk


def mesh_rank(rnk_n, rnk_pm):
    # util/util.c:253-260: 3-d FFT on a 3-d mesh is remapped to 2-d
    return 2 if (rnk_n == 3 and rnk_pm == 3) else rnk_pm


def local_size(n, ni, no, howmany, iblock, oblock, rnk_pm, tf, flags):
    """PX(local_size_partrafo) on a one-process mesh of dimension rnk_pm.
    Returns (mem, local_ni, local_i_start, local_no, local_o_start)."""
    rnk_n = len(n)
    transp = flags & (TRANSPOSED_IN | TRANSPOSED_OUT)
    if transp == (TRANSPOSED_IN | TRANSPOSED_OUT):
        # the reference writes none of the output arrays in this case
        raise ValueError("TRANSPOSED_IN|TRANSPOSED_OUT leaves the local sizes unset")
    r = mesh_rank(rnk_n, rnk_pm)
    np_pm = [1] * r
    coords = [0] * r
    iblk, mblk, oblk = evaluate_blocks(ni, no, iblock, oblock, r, np_pm, tf, transp)
    to, ti = init_param_size_and_trafo_flags(n, ni, no, r, tf, transp, None)
    dummy = None
    mem = 1
    res_to = res_ti = None
    if not transp & TRANSPOSED_IN:
        m, a, b, c, d = local_size_partrafo_transposed(to[0], to[1], to[2], howmany, iblk, mblk, r,
                                                        coords, TRANSPOSED_OUT, to[3])
        mem = max(mem, m)
        res_to = (a, b, c, d)
    if not transp & TRANSPOSED_OUT:
        m, a, b, c, d = local_size_partrafo_transposed(ti[0], ti[1], ti[2], howmany, mblk, oblk, r,
                                                        coords, TRANSPOSED_IN, ti[3])
        mem = max(mem, m)
        res_ti = (a, b, c, d)
    # init_param_local_size: which step owns the user's arrays
    if transp & TRANSPOSED_OUT:
        lni, lis, lno, los = res_to
    elif transp & TRANSPOSED_IN:
        lni, lis, lno, los = res_ti
    else:
        lni, lis = res_to[0], res_to[1]
        lno, los = res_ti[2], res_ti[3]
    lis, los = list(lis), list(los)
    if flags & SHIFTED_IN:
        for t in range(rnk_n):
            lis[t] -= ni[t] // 2
    if flags & SHIFTED_OUT:
        for t in range(rnk_n):
            los[t] -= no[t] // 2
    return mem, list(lni), lis, list(lno), los


def local_block(ni, no, iblock, oblock, rnk_pm, pid, tf, flags):
    """PX(local_block_partrafo) for process `pid` (only 0 exists)."""
    assert pid == 0
    rnk_n = len(ni)
    transp = flags & (TRANSPOSED_IN | TRANSPOSED_OUT)
    r = mesh_rank(rnk_n, rnk_pm)
    np_pm = [1] * r
    coords = [0] * r
    iblk, mblk, oblk = evaluate_blocks(ni, no, iblock, oblock, r, np_pm, tf, transp)
    to, ti = init_param_size_and_trafo_flags(ni, ni, no, r, tf, transp, None)
    res_to = res_ti = None
    if not transp & TRANSPOSED_IN:
        a, b, c, d = local_size_transposed(to[1], to[2], iblk, mblk, r, coords, to[3][r],
                                           TRANSPOSED_OUT)
        overwrite_real_counts(to[1], to[2], to[3][r], a, c)
        res_to = (a, b, c, d)
    if not transp & TRANSPOSED_OUT:
        a, b, c, d = local_size_transposed(ti[1], ti[2], mblk, oblk, r, coords, ti[3][r],
                                           TRANSPOSED_IN)
        overwrite_real_counts(ti[1], ti[2], ti[3][r], a, c)
        res_ti = (a, b, c, d)
    if transp & TRANSPOSED_OUT:
        lni, lis, lno, los = res_to
    elif transp & TRANSPOSED_IN:
        lni, lis, lno, los = res_ti
    else:
        lni, lis = res_to[0], res_to[1]
        lno, los = res_ti[2], res_ti[3]
    lis, los = list(lis), list(los)
    if flags & SHIFTED_IN:
        for t in range(rnk_n):
            lis[t] -= ni[t] // 2
    if flags & SHIFTED_OUT:
        for t in range(rnk_n):
            los[t] -= no[t] // 2
    return list(lni), lis, list(lno), los


def plan_is_null(n, rnk_pm, tf, flags):
    """kernel/partrafo.c:337-373: True where PX(plan_partrafo) returns NULL."""
    rnk_n = len(n)
    transp = flags & (TRANSPOSED_IN | TRANSPOSED_OUT)
    if flags & (SHIFTED_IN | SHIFTED_OUT):
        if any(x % 2 for x in n):
            return True
    if (transp & TRANSPOSED_OUT) and (transp & TRANSPOSED_IN):
        return True
    if (tf & T_R2C) and (transp & TRANSPOSED_IN):
        return True
    if (tf & T_C2R) and (transp & TRANSPOSED_OUT):
        return True
    if rnk_n < rnk_pm:
        return True
    if rnk_n == rnk_pm and rnk_n != 3:
        return True
    return False


def transposed_order(rnk_n, r):
    """Memory order of the transposed decomposition
    n1 x .. x n_r x n0 x n_{r+1} x .. (doc/tutorial.tex:319-321)."""
    return list(range(1, r + 1)) + [0] + list(range(r + 1, rnk_n))
