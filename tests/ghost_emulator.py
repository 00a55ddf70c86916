"""Host-side interpreter of the device ghost-fill program (test helper).

Executes the op list produced by Forest.ghost_program on numpy arrays with
the same per-op semantics as the ghost_phase kernel, so the op builder (host
logic) can be checked against the oracle without a GPU.
"""
import numpy as np

COPY_FACE, COPY_CORNER, AVG_FACE, AVG_CORNER, INTERP_FACE, INTERP_CORNER, PHYSBC = range(7)


def _slope(sl, sr):
    sc = (sl + sr) / 2.0
    mn = min(min(2 * abs(sl), 2 * abs(sr)), abs(sc))
    return mn * max(0.0, np.copysign(1.0, sl * sr)) * np.copysign(1.0, sc)


def run(q, ops, phase_start, mx, my, mbc):
    """q: (npatch, meqn, my+2mbc, mx+2mbc), modified in place."""
    meqn = q.shape[1]
    o = mbc - 1  # array index of cell i is i + o

    def Q(p, i, j, m):
        return q[p, m, j + o, i + o]

    def S(p, i, j, m, v):
        q[p, m, j + o, i + o] = v

    for ph in range(6):
        for k in range(phase_start[ph], phase_start[ph + 1]):
            t, dst, src, which, ti, tj, bc, _ = (int(v) for v in ops[k])
            for m in range(meqn):
                if t == COPY_FACE:
                    if which < 2:
                        ir = range(1 - mbc, 1) if which == 0 else range(mx + 1, mx + mbc + 1)
                        jr = range(1, my + 1)
                    else:
                        ir = range(1, mx + 1)
                        jr = range(1 - mbc, 1) if which == 2 else range(my + 1, my + mbc + 1)
                    for j in jr:
                        for i in ir:
                            S(dst, i, j, m, Q(src, i + ti, j + tj, m))
                elif t == COPY_CORNER or t == AVG_CORNER:
                    for ibc in range(1, mbc + 1):
                        for jbc in range(1, mbc + 1):
                            i = mx + ibc if which & 1 else 1 - ibc
                            j = my + jbc if which & 2 else 1 - jbc
                            if t == COPY_CORNER:
                                S(dst, i, j, m, Q(src, i + ti, j + tj, m))
                            else:
                                fi, fj = 2 * i - 1 + ti, 2 * j - 1 + tj
                                s = 0.0
                                for a, b in ((0, 0), (1, 0), (0, 1), (1, 1)):
                                    s = s + Q(src, fi + a, fj + b, m)
                                S(dst, i, j, m, s / 4.0)
                elif t == AVG_FACE:
                    length = my if which < 2 else mx
                    for l in range(length):
                        for b in range(1, mbc + 1):
                            if which < 2:
                                ic, jc = (1 - b if which == 0 else mx + b), 1 + l
                            else:
                                ic, jc = 1 + l, (1 - b if which == 2 else my + b)
                            fi, fj = 2 * ic - 1 + ti, 2 * jc - 1 + tj
                            if fi < 1 or fi + 1 > mx or fj < 1 or fj + 1 > my:
                                continue
                            s = 0.0
                            for a, bb in ((0, 0), (1, 0), (0, 1), (1, 1)):
                                s = s + Q(src, fi + a, fj + bb, m)
                            S(dst, ic, jc, m, s / 4.0)
                elif t in (INTERP_FACE, INTERP_CORNER):
                    hb = mbc // 2
                    cells = []
                    if t == INTERP_FACE:
                        for l in range(mx):
                            for b in range(1, hb + 1):
                                if which < 2:
                                    cells.append(((b if which == 0 else mx - b + 1), 1 + l))
                                else:
                                    cells.append((1 + l, (b if which == 2 else my - b + 1)))
                    else:
                        for ibc in range(1, hb + 1):
                            for jbc in range(1, hb + 1):
                                cells.append((mx - ibc + 1 if which & 1 else ibc,
                                              my - jbc + 1 if which & 2 else jbc))
                    for ic, jc in cells:
                        fi, fj = 2 * ic - 1 + ti, 2 * jc - 1 + tj
                        if fi < 1 - mbc or fi + 1 > mx + mbc or fj < 1 - mbc or fj + 1 > my + mbc:
                            continue
                        c = Q(src, ic, jc, m)
                        gx = _slope(c - Q(src, ic - 1, jc, m), Q(src, ic + 1, jc, m) - c)
                        gy = _slope(c - Q(src, ic, jc - 1, m), Q(src, ic, jc + 1, m) - c)
                        for a, b, sx, sy in ((0, 0, -0.25, -0.25), (1, 0, 0.25, -0.25),
                                             (0, 1, -0.25, 0.25), (1, 1, 0.25, 0.25)):
                            S(dst, fi + a, fj + b, m, c + gx * sx + gy * sy)
            if t == PHYSBC:
                for face in range(4):
                    mth = ((bc >> (8 * face)) & 0xFF)
                    mth = mth - 256 if mth > 127 else mth
                    if mth not in (1, 3):
                        continue
                    for m in range(meqn):
                        for b in range(1, mbc + 1):
                            n_l = my + 2 * mbc if face < 2 else mx + 2 * mbc
                            for l in range(n_l):
                                if face < 2:
                                    jd = js = 1 - mbc + l
                                    if face == 0:
                                        idd, is_ = 1 - b, (1 if mth == 1 else b)
                                    else:
                                        idd, is_ = mx + b, (mx if mth == 1 else mx + 1 - b)
                                else:
                                    idd = is_ = 1 - mbc + l
                                    if face == 2:
                                        jd, js = 1 - b, (1 if mth == 1 else b)
                                    else:
                                        jd, js = my + b, (my if mth == 1 else my + 1 - b)
                                mn = 1 if face < 2 else 2
                                v = Q(dst, is_, js, m)
                                S(dst, idd, jd, m, -v if (mth == 3 and m == mn) else v)
    return q


def _bc_code(bc, face):
    mth = (bc >> (8 * face)) & 0xFF
    return mth - 256 if mth > 127 else mth


def face_table(ops, npatch):
    """What fills each face ghost strip, from the ops (the fused launch's face
    table): ('same', nb, ti, tj) / ('half', [(nb, ti, tj), ...]) /
    ('double',) / ('phys', mth) / None."""
    rel = [[None] * 4 for _ in range(npatch)]
    for op in ops:
        t, dst, src, which, ti, tj, bc, _ = (int(v) for v in op)
        if t == COPY_FACE:
            rel[dst][which] = ('same', src, ti, tj)
        elif t == AVG_FACE:
            if rel[dst][which] is None or rel[dst][which][0] != 'half':
                rel[dst][which] = ('half', [])
            rel[dst][which][1].append((src, ti, tj))
        elif t == INTERP_FACE:
            rel[dst][which ^ 1] = ('double',)
        elif t == PHYSBC:
            for f in range(4):
                mth = _bc_code(bc, f)
                if mth in (1, 3):
                    rel[dst][f] = ('phys', mth)
    return rel


def run_flat(q, ops, phase_start, mx, my, mbc):
    """The semantics of the fused launch (ghost_flat) with every op: each ghost
    cell computed from interior cells only -- copies and averages read interior
    cells; interpolation reads the coarse patch's face strips recomputed from the
    face table; the physical BC (once per patch) reads its sources the same way,
    an interpolated strip included, and writes each cell once (a corner of two
    physical faces takes the y-face's value)."""
    npatch, meqn = q.shape[0], q.shape[1]
    src_q = q.copy()
    o = mbc - 1
    rel = face_table(ops, npatch)
    interp_src = {}
    for op in ops:
        t, dst, src, which, ti, tj, bc, _ = (int(v) for v in op)
        if t == INTERP_FACE:
            interp_src[(dst, which ^ 1)] = (src, ti, tj)

    def Qs(p, i, j, m):
        return src_q[p, m, j + o, i + o]

    def simple(c, i, j, m):
        k = 0 if i < 1 else 1 if i > mx else 2 if j < 1 else 3 if j > my else -1
        if k < 0:
            return Qs(c, i, j, m)
        r = rel[c][k]
        assert r is not None and r[0] in ('same', 'half', 'phys'), (c, k, r)
        if r[0] == 'same':
            return Qs(r[1], i + r[2], j + r[3], m)
        if r[0] == 'half':
            for nb, ti, tj in r[1]:
                fi, fj = 2 * i - 1 + ti, 2 * j - 1 + tj
                if 1 <= fi and fi + 1 <= mx and 1 <= fj and fj + 1 <= my:
                    s = 0.0
                    for a, b in ((0, 0), (1, 0), (0, 1), (1, 1)):
                        s = s + Qs(nb, fi + a, fj + b, m)
                    return s / 4.0
            raise AssertionError("no child covers the cell")
        mth = r[1]
        i2, j2, mn = i, j, (1 if k < 2 else 2)
        if k == 0:
            i2 = 1 if mth == 1 else 1 - i
        elif k == 1:
            i2 = mx if mth == 1 else 2 * mx + 1 - i
        elif k == 2:
            j2 = 1 if mth == 1 else 1 - j
        else:
            j2 = my if mth == 1 else 2 * my + 1 - j
        v = Qs(c, i2, j2, m)
        return -v if (mth == 3 and m == mn) else v

    def interp_child(c, ic, jc, m, a, b):
        cc = simple(c, ic, jc, m)
        gx = _slope(cc - simple(c, ic - 1, jc, m), simple(c, ic + 1, jc, m) - cc)
        gy = _slope(cc - simple(c, ic, jc - 1, m), simple(c, ic, jc + 1, m) - cc)
        return cc + gx * (0.25 if a else -0.25) + gy * (0.25 if b else -0.25)

    def cell(p, i, j, m):
        k = 0 if i < 1 else 1 if i > mx else 2 if j < 1 else 3 if j > my else -1
        if k >= 0 and rel[p][k] is not None and rel[p][k][0] == 'double':
            c, ti, tj = interp_src[(p, k)]
            u, v = i - ti + 1, j - tj + 1
            return interp_child(c, u >> 1, v >> 1, m, u & 1, v & 1)
        return simple(p, i, j, m)

    bc_last = {}
    for ph in range(6):
        for k in range(phase_start[ph], phase_start[ph + 1]):
            if int(ops[k][0]) == PHYSBC:
                bc_last[int(ops[k][1])] = k
    for k, op in enumerate(ops):
        t, dst, src, which, ti, tj, bc, _ = (int(v) for v in op)
        if t in (COPY_FACE, COPY_CORNER, AVG_FACE, AVG_CORNER):
            tmp = src_q.copy()
            run(tmp, np.array([op], dtype=np.int32), [0, 1, 1, 1, 1, 1, 1], mx, my, mbc)
            changed = tmp[dst] != src_q[dst]
            # the op's cells: wherever the sequential op wrote (compare with NaN marks)
            mark = src_q.copy()
            mark[dst] = np.nan
            run(mark, np.array([op], dtype=np.int32), [0, 1, 1, 1, 1, 1, 1], mx, my, mbc)
            w = ~np.isnan(mark[dst])
            q[dst][w] = tmp[dst][w]
        elif t in (INTERP_FACE, INTERP_CORNER):
            hb = mbc // 2
            cells = []
            if t == INTERP_FACE:
                for l in range(mx):
                    for b in range(1, hb + 1):
                        if which < 2:
                            cells.append(((b if which == 0 else mx - b + 1), 1 + l))
                        else:
                            cells.append((1 + l, (b if which == 2 else my - b + 1)))
            else:
                for ibc in range(1, hb + 1):
                    for jbc in range(1, hb + 1):
                        cells.append((mx - ibc + 1 if which & 1 else ibc,
                                      my - jbc + 1 if which & 2 else jbc))
            for m in range(meqn):
                for ic, jc in cells:
                    fi, fj = 2 * ic - 1 + ti, 2 * jc - 1 + tj
                    if fi < 1 - mbc or fi + 1 > mx + mbc or fj < 1 - mbc or fj + 1 > my + mbc:
                        continue
                    for a in (0, 1):
                        for b in (0, 1):
                            q[dst, m, fj + b + o, fi + a + o] = interp_child(src, ic, jc, m, a, b)
        elif t == PHYSBC and bc_last.get(dst) == k:
            phys = [_bc_code(bc, f) in (1, 3) for f in range(4)]
            for face in range(4):
                if not phys[face]:
                    continue
                mth = _bc_code(bc, face)
                for m in range(meqn):
                    for b in range(1, mbc + 1):
                        n_l = my + 2 * mbc if face < 2 else mx + 2 * mbc
                        for l in range(n_l):
                            if face < 2:
                                jd = js = 1 - mbc + l
                                if (jd < 1 and phys[2]) or (jd > my and phys[3]):
                                    continue
                                if face == 0:
                                    idd, is_ = 1 - b, (1 if mth == 1 else b)
                                else:
                                    idd, is_ = mx + b, (mx if mth == 1 else mx + 1 - b)
                            else:
                                idd = is_ = 1 - mbc + l
                                if face == 2:
                                    jd, js = 1 - b, (1 if mth == 1 else b)
                                else:
                                    jd, js = my + b, (my if mth == 1 else my + 1 - b)
                            mn = 1 if face < 2 else 2
                            v = cell(dst, is_, js, m)
                            q[dst, m, jd + o, idd + o] = -v if (mth == 3 and m == mn) else v
    return q
