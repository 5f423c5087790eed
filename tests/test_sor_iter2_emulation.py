"""CPU emulation of the two-rows-per-thread SOR iteration (k_sor_iter2 in
mpi-les_amd/csrc/k_press.hip): its block/tile/row/lane index arithmetic,
LDS planes, register rings, store redirection and per-cell FP32 arithmetic,
statement by statement, checked against a plain red-black iteration with the
ghost rules of press.f95:89-122 / boundp.f95:7-78 (red from the previous
black iterate, then black from the new red), bit for bit, for tilings with
partial tiles, tiny jm and short k chunks, with 16 x 64, 32 x 64 and
32 x 32 tiles (rows x pairs with the ring; in the last, a wave holds two
half-wave rows).  It also checks that every buffer
and LDS index the kernel forms stays inside its allocation, that every
interior cell is stored exactly once by its owner, and that the redirected
stores land only in the row padding or the scratch plane.  The GPU tests
check the kernel itself against the colour sweeps; this pins the index
arithmetic on the CPU."""
import numpy as np
import pytest

FX = 64  # threads per wave (blockDim.x)
f32 = np.float32


def jwrap(j, jm):
    return (j - 1) % jm + 1


def xcd(b, nb):
    x, l = b % 8, b // 8
    base, rem = nb // 8, nb % 8
    return x * base + (x if x < rem else rem) + l


class Err(Exception):
    pass


def run(im, jm, km, kc, ty=16, seed=0, omega=1.0, remap=1, tw=FX):
    assert jm % 2 == 0
    assert tw == FX or (tw == FX // 2 and ty == 32)
    FY = ty  # tile rows with the ring
    FTY = FY - 4
    W = tw  # tile width in pairs with the ring
    FTXW = W - 4
    NW = FY * W // (2 * FX)  # waves of the block
    PL = FY * W
    rng = np.random.default_rng(seed)
    nt = (im + 1) // 2
    cp = (nt + 1 + 31) // 32 * 32
    cpl = cp * jm
    nplanes = 8 * km + 2
    mem = rng.uniform(-1, 1, nplanes * cpl).astype(f32)  # whole allocation, garbage everywhere
    # array a (0..7) level L (1..km) at plane 1 + a*km + (L-1)
    def plane(a, L):
        return 1 + a * km + (L - 1)
    # arrays: 0 p[0][0] (par0 black), 1 p[0][1] (par0 red), 2 p[1][0], 3 p[1][1], 4 rhs black, 5 rhs red, 6 cn1 b, 7 cn1 r
    prv, cur = 0, 1  # iter = 1 -> prv = 0, cur = 1
    A_Bo, A_Ro, A_Bn, A_Rn, A_rB, A_rR = 2 * prv + 0, 2 * prv + 1, 2 * cur + 0, 2 * cur + 1, 4, 5
    # metrics
    c2l = np.zeros(im + 4, f32); c2s = np.zeros(im + 4, f32); ax = np.zeros(im + 4, f32)
    c3l = np.zeros(jm + 4, f32); c3s = np.zeros(jm + 4, f32); by = np.zeros(jm + 4, f32)
    c4l = np.zeros(km + 4, f32); c4s = np.zeros(km + 4, f32); cz = np.zeros(km + 4, f32)
    for a in (c2l, c2s, c3l, c3s, c4l, c4s):
        a[:] = rng.uniform(0.05, 0.3, a.size).astype(f32)
    for a in (ax, by, cz):
        a[:] = rng.uniform(0.1, 0.5, a.size).astype(f32)
    M1 = lambda a, x: a[x + 1]  # noqa: E731

    def split_get(a, t, j, L):
        return mem[plane(a, L) * cpl + (j - 1) * cp + t]

    # ---------------- reference: red then black over the interior
    def pos_red(j, k):
        return (j + k) & 1

    old = {}
    for colour, arr in (("R", A_Ro), ("B", A_Bo)):
        for k in range(1, km + 1):
            for j in range(1, jm + 1):
                for t in range(nt):
                    pr = pos_red(j, k)
                    pos = pr if colour == "R" else 1 - pr
                    i = 2 * t + 1 + pos
                    if i <= im:
                        old[(i, j, k)] = split_get(arr, t, j, k)
    rhs = {}
    for colour, arr in (("R", A_rR), ("B", A_rB)):
        for k in range(1, km + 1):
            for j in range(1, jm + 1):
                for t in range(nt):
                    pr = pos_red(j, k)
                    pos = pr if colour == "R" else 1 - pr
                    i = 2 * t + 1 + pos
                    if i <= im:
                        rhs[(i, j, k)] = split_get(arr, t, j, k)
    new = dict(old)
    om = f32(omega)

    def upd(i, j, k, src):
        own = src[(i, j, k)] if src is old else old[(i, j, k)]
        nb = src
        pim = own if i == 1 else nb[(i - 1, j, k)]
        pip = own if i == im else nb[(i + 1, j, k)]
        pjm = nb[(i, jwrap(j - 1, jm), k)]
        pjp = nb[(i, jwrap(j + 1, jm), k)]
        pkm = own if k == 1 else nb[(i, j, k - 1)]
        pkp = own if k == km else nb[(i, j, k + 1)]
        cn = f32(1.0) / f32(f32(M1(ax, i) + M1(by, j)) + M1(cz, k))
        s = f32(M1(c2l, i) * pip)
        s = f32(s + f32(M1(c2s, i) * pim))
        s = f32(s + f32(M1(c3l, j) * pjp))
        s = f32(s + f32(M1(c3s, j) * pjm))
        s = f32(s + f32(M1(c4l, k) * pkp))
        s = f32(s + f32(M1(c4s, k) * pkm))
        s = f32(s - rhs[(i, j, k)])
        r = f32(om * f32(f32(cn * s) - own))
        return f32(own + r), r

    ref_sum = 0.0
    for k in range(1, km + 1):
        for j in range(1, jm + 1):
            for i in range(1, im + 1):
                if (i + j + k) & 1:  # red
                    v, r = upd(i, j, k, old)
                    new[(i, j, k)] = v
                    ref_sum += float(r) * float(r)
    newb = dict(new)
    for k in range(1, km + 1):
        for j in range(1, jm + 1):
            for i in range(1, im + 1):
                if not (i + j + k) & 1:  # black: neighbours are new red
                    own = old[(i, j, k)]
                    pim = own if i == 1 else new[(i - 1, j, k)]
                    pip = own if i == im else new[(i + 1, j, k)]
                    pjm = new[(i, jwrap(j - 1, jm), k)]
                    pjp = new[(i, jwrap(j + 1, jm), k)]
                    pkm = own if k == 1 else new[(i, j, k - 1)]
                    pkp = own if k == km else new[(i, j, k + 1)]
                    cn = f32(1.0) / f32(f32(M1(ax, i) + M1(by, j)) + M1(cz, k))
                    s = f32(M1(c2l, i) * pip)
                    s = f32(s + f32(M1(c2s, i) * pim))
                    s = f32(s + f32(M1(c3l, j) * pjp))
                    s = f32(s + f32(M1(c3s, j) * pjm))
                    s = f32(s + f32(M1(c4l, k) * pkp))
                    s = f32(s + f32(M1(c4s, k) * pkm))
                    s = f32(s - rhs[(i, j, k)])
                    r = f32(om * f32(f32(cn * s) - own))
                    newb[(i, j, k)] = f32(own + r)
                    ref_sum += float(r) * float(r)

    # ---------------- emulation of the kernel
    whole = nt <= W  # one tile spans the width: no ring lanes (the i walls are mirrored)
    ntt, ntj = (1 if whole else (nt + FTXW - 1) // FTXW), (jm + FTY - 1) // FTY
    nz = (km + kc - 1) // kc
    nparts = ntt * ntj * nz
    cplb = 4 * cpl
    nbytes = nplanes * cplb
    off0 = lambda a: (plane(a, 1) - 1) * cplb  # noqa: E731  level-0 byte offset
    oRo, oBo, oRn, oBn, orR, orB = (off0(a) for a in (A_Ro, A_Bo, A_Rn, A_Bn, A_rR, A_rB))
    written = {}  # (flat index) -> list of (block, value)
    emu_sum = 0.0

    def load(voff, soff):
        addr = voff + soff
        if np.any(addr < 0) or np.any(addr + 4 > nbytes) or np.any(addr % 4):
            raise Err(f"load out of range {addr.min()}..{addr.max()} of {nbytes}")
        return mem[addr // 4].copy()

    out = mem.copy()

    def store(val, voff, soff, bid, real_mask):
        addr = voff + soff
        if np.any(addr < 0) or np.any(addr + 4 > nbytes):
            raise Err("store out of range")
        idx = (addr // 4).ravel()
        for q, v, rm in zip(idx, val.ravel(), real_mask.ravel()):
            written.setdefault(int(q), []).append((bid, float(v), bool(rm)))
            out[q] = v

    # thread (wave w, x = threadIdx.x): lane = pair of the tile row, wv = its pair of rows
    tx = np.arange(FX)[None, :].repeat(NW, 0)
    ww = np.arange(NW)[:, None].repeat(FX, 1)
    lane = tx % W
    wv = ww * (FX // W) + tx // W
    for bid in range(nparts):
        lb = xcd(bid, nparts) if remap else bid
        bt, bj, bk = lb % ntt, (lb // ntt) % ntj, lb // (ntt * ntj)
        t = lane if whole else bt * FTXW - 2 + lane
        tin = (t >= 0) & (t < nt)
        rA = 2 * wv
        rB = rA + 1
        jlA = 1 + bj * FTY - 2 + rA
        jlB = jlA + 1
        jA = jwrap(jlA, jm)
        jB = jwrap(jlB, jm)
        assert np.all(jA % 2 == 1)
        k0 = 1 + bk * kc
        k1 = min(km, k0 + kc - 1)
        krs = max(1, k0 - 1)
        kre = k1 + 1
        tq = np.where(tin, t, 0)
        voA = 4 * (tq + (jA - 1) * cp)
        voB = 4 * (tq + (jB - 1) * cp)
        trA = 4 * (nt + (jA - 1) * cp)
        trB = 4 * (nt + (jB - 1) * cp)
        lane_tile = tin & (whole | ((lane >= 2) & (lane <= W - 3)))
        tileA = (rA >= 2) & (rA <= FY - 3) & (jlA <= jm)
        tileB = (rB >= 2) & (rB <= FY - 3) & (jlB <= jm)
        c2lp = [np.where(tin, c2l[np.clip(2 * t + 1 + p, -1, im + 2) + 1], 0).astype(f32) for p in (0, 1)]
        c2sp = [np.where(tin, c2s[np.clip(2 * t + 1 + p, -1, im + 2) + 1], 0).astype(f32) for p in (0, 1)]
        axp = [np.where(tin, ax[np.clip(2 * t + 1 + p, -1, im + 2) + 1], 0).astype(f32) for p in (0, 1)]
        c3lA, c3sA, c3lB, c3sB = c3l[jA + 1], c3s[jA + 1], c3l[jB + 1], c3s[jB + 1]
        byA, byB = by[jA + 1], by[jB + 1]
        sA = [f32(1) * (axp[p] + byA).astype(f32) for p in (0, 1)]
        sB = [(axp[p] + byB).astype(f32) for p in (0, 1)]
        okA = [lane_tile & tileA & (2 * t + 1 + p <= im) for p in (0, 1)]
        okB = [lane_tile & tileB & (2 * t + 1 + p <= im) for p in (0, 1)]
        svA = [np.where(okA[p], voA, trA) for p in (0, 1)]
        svB = [np.where(okB[p], voB, trB) for p in (0, 1)]
        last = [(2 * t + 1 + p == im) for p in (0, 1)]
        first = (t == 0)
        ownA = rA * W + lane
        bsA = np.clip(ownA - 1, 0, PL - 3)
        bsB = np.clip(ownA + W - 1, 0, PL - 3)
        upA = np.where(rA == 0, rA, rA - 1) * W + lane
        dnB = np.where(rB == FY - 1, rB, rB + 1) * W + lane
        nsteps = kre - krs + 1
        lvm = np.zeros((nsteps, 3), f32)
        for x in range(nsteps):
            k = krs + x
            if k <= km:
                lvm[x] = (M1(c4l, k), M1(c4s, k), M1(cz, k))
        lbq = min(kre + 1, km + 1)

        def fill(L):
            lr = min(L, kre) * cplb
            lk = min(L - 1, kre - 1) * cplb
            return dict(roA=load(voA, oRo + lr), roB=load(voB, oRo + lr), rrA=load(voA, orR + lr),
                        rrB=load(voB, orR + lr), rbA=load(voA, orB + lk), rbB=load(voB, orB + lk))

        def loadbq(L):
            # level km+1 (the boundp2 ghost of level km) reads the red iterate of level km
            Lc = min(L, lbq)
            o = oRo + km * cplb if Lc == km + 1 else oBo + Lc * cplb
            return load(voA, o), load(voB, o)

        sl = [fill(krs), fill(krs + 1), fill(krs + 2)]
        bqA = [None] * 6
        bqB = [None] * 6
        for q in range(5):
            bqA[q], bqB[q] = loadbq(krs - 1 + q)
        LB = np.full(3 * PL, np.nan, f32)
        LR = np.full(2 * PL, np.nan, f32)
        LB[ownA] = bqA[1]
        LB[ownA + W] = bqB[1]
        rdA = [np.zeros_like(voA, dtype=f32)] * 3
        rdB = [np.zeros_like(voA, dtype=f32)] * 3
        if krs == 1:  # level 0 (the boundp2 ghost of level 1): the cells themselves
            bqA[0], bqB[0] = sl[0]["roA"], sl[0]["roB"]
            rdA = list(rdA); rdB = list(rdB)
            rdA[2], rdB[2] = bqA[1], bqB[1]
        mx = [f32(0), f32(0)]
        my = [f32(0), f32(0)]
        mz = [f32(0), f32(0)]
        D0 = (1 + krs) & 1
        acc = np.zeros(voA.shape)

        def lds_read(arr, idx, size):
            if np.any(idx < 0) or np.any(idx >= size):
                raise Err("LDS out of range")
            v = arr[idx]
            return v

        def cellupd(own, lo, hi, pjm, pjp, pkm, pkp, p, c3l_, c3s_, s_, mxx, myy, mzz, rr, first_, last_, kf, kl):
            pim = np.where(first_, own, lo) if p == 0 else lo
            pip = np.where(last_, own, hi)
            pkm = own if kf else pkm
            pkp = own if kl else pkp
            cn = (f32(1) / (s_ + mzz).astype(f32)).astype(f32)
            s = (c2lp[p] * pip).astype(f32)
            s = (s + (c2sp[p] * pim).astype(f32)).astype(f32)
            s = (s + (c3l_ * pjp).astype(f32)).astype(f32)
            s = (s + (c3s_ * pjm).astype(f32)).astype(f32)
            s = (s + (mxx * pkp).astype(f32)).astype(f32)
            s = (s + (myy * pkm).astype(f32)).astype(f32)
            s = (s - rr).astype(f32)
            r = (f32(omega) * ((cn * s).astype(f32) - own).astype(f32)).astype(f32)
            return (own + r).astype(f32), r

        for U_abs, kr in enumerate(range(krs, kre + 1)):
            U = U_abs % 6
            P = U & 1
            pa = D0 ^ P
            pb = pa ^ 1
            kb = kr - 1
            LB[((U + 1) % 3) * PL + ownA] = bqA[(U + 2) % 6]
            LB[((U + 1) % 3) * PL + ownA + W] = bqB[(U + 2) % 6]
            if not (0 <= kr - krs < nsteps):
                raise Err("lvm index")
            mx[P], my[P], mz[P] = lvm[kr - krs]
            kinR = k0 <= kr <= k1
            kinB = k0 <= kb <= k1
            soR = oRn + kr * cplb if kinR else 0
            soB = oBn + kb * cplb if kinB else 0
            kf, kl, bf, bl = kr == 1, kr == km, kb == 1, kb == km
            q = sl[U % 3]
            pbase = (U % 3) * PL
            lo = lds_read(LB, pbase + bsA + pa, 3 * PL); hi = lds_read(LB, pbase + bsA + pa + 1, 3 * PL)
            redA, rA_ = cellupd(q["roA"], lo, hi, lds_read(LB, pbase + upA, 3 * PL), bqB[(U + 1) % 6],
                                bqA[U % 6], bqA[(U + 2) % 6], pa, c3lA, c3sA, sA[pa], mx[P], my[P], mz[P],
                                q["rrA"], first, last[pa], False, False)
            store(redA, svA[pa], soR, bid, okA[pa] & kinR)
            acc += np.where(okA[pa] & kinR, rA_.astype(float) ** 2, 0)
            lo = lds_read(LB, pbase + bsB + pb, 3 * PL); hi = lds_read(LB, pbase + bsB + pb + 1, 3 * PL)
            redB, rB_ = cellupd(q["roB"], lo, hi, bqA[(U + 1) % 6], lds_read(LB, pbase + dnB, 3 * PL),
                                bqB[U % 6], bqB[(U + 2) % 6], pb, c3lB, c3sB, sB[pb], mx[P], my[P], mz[P],
                                q["rrB"], first, last[pb], False, False)
            store(redB, svB[pb], soR, bid, okB[pb] & kinR)
            acc += np.where(okB[pb] & kinR, rB_.astype(float) ** 2, 0)
            LR[P * PL + ownA] = redA
            LR[P * PL + ownA + W] = redB
            rdA = list(rdA); rdB = list(rdB)
            rdA[U % 3] = redA
            rdB[U % 3] = redB
            rbase = (P ^ 1) * PL
            lo = lds_read(LR, rbase + bsA + pa, 2 * PL); hi = lds_read(LR, rbase + bsA + pa + 1, 2 * PL)
            ownb = bqA[U % 6]
            bA, bA_ = cellupd(ownb, lo, hi, lds_read(LR, rbase + upA, 2 * PL), rdB[(U + 2) % 3],
                              rdA[(U + 1) % 3], redA, pa, c3lA, c3sA, sA[pa], mx[P ^ 1], my[P ^ 1], mz[P ^ 1],
                              q["rbA"], first, last[pa], False, bl)
            store(bA, svA[pa], soB, bid, okA[pa] & kinB)
            acc += np.where(okA[pa] & kinB, bA_.astype(float) ** 2, 0)
            lo = lds_read(LR, rbase + bsB + pb, 2 * PL); hi = lds_read(LR, rbase + bsB + pb + 1, 2 * PL)
            ownb = bqB[U % 6]
            bB, bB_ = cellupd(ownb, lo, hi, rdA[(U + 2) % 3], lds_read(LR, rbase + dnB, 2 * PL),
                              rdB[(U + 1) % 3], redB, pb, c3lB, c3sB, sB[pb], mx[P ^ 1], my[P ^ 1], mz[P ^ 1],
                              q["rbB"], first, last[pb], False, bl)
            store(bB, svB[pb], soB, bid, okB[pb] & kinB)
            acc += np.where(okB[pb] & kinB, bB_.astype(float) ** 2, 0)
            sl[U % 3] = fill(kr + 3)
            bqA[(U + 5) % 6], bqB[(U + 5) % 6] = loadbq(kr + 4)
        emu_sum += acc.sum()

    # ---------------- compare
    bad = 0
    for (i, j, k), v in newb.items():
        red = (i + j + k) & 1
        t = (i - 1) // 2
        arr = A_Rn if red else A_Bn
        qidx = plane(arr, k) * cpl + (j - 1) * cp + t
        ws = written.get(qidx)
        if not ws:
            bad += 1
            if bad < 5:
                print("unwritten", (i, j, k), "red" if red else "black")
            continue
        reals = [x for x in ws if x[2]]
        if len(reals) != 1:
            bad += 1
            if bad < 5:
                print("owned writes", len(reals), (i, j, k))
        vals = set(np.float32(x[1]).tobytes() for x in ws)
        if len(vals) != 1 or np.float32(ws[-1][1]) != v:
            bad += 1
            if bad < 5:
                print("mismatch", (i, j, k), "red" if red else "black", [x[1] for x in ws], float(v))
    # every non-owned write must land in padding (pair >= nt) or plane 0, or equal the owned value of that cell
    for qidx, ws in written.items():
        pl, rem = divmod(qidx, cpl)
        jrow, tt = divmod(rem, cp)
        if pl == 0 or tt >= nt:
            if any(x[2] for x in ws):
                bad += 1
                print("owned write into scratch", qidx)
            continue
        if not any(x[2] for x in ws):
            bad += 1
            if bad < 10:
                print("non-owned write into a real cell", pl, jrow + 1, tt)
    rel = abs(emu_sum - ref_sum) / max(ref_sum, 1e-30)
    return bad, rel, nparts


CASES = [(20, 14, 9, 9), (20, 14, 9, 3), (21, 14, 9, 4), (130, 26, 12, 5), (7, 2, 2, 1), (9, 4, 3, 2),
         (125, 30, 20, 7), (64, 16, 13, 13), (3, 2, 5, 5), (121, 12, 8, 6), (70, 60, 11, 4),
         (128, 18, 6, 3), (127, 10, 5, 5), (129, 8, 4, 2)]  # 64 pairs: one whole-width tile; 65: two


@pytest.mark.parametrize("ty,tw", [(16, FX), (32, FX), (32, FX // 2)])
@pytest.mark.parametrize("im,jm,km,kc", CASES)
def test_sor_iter2_tiling_emulation(im, jm, km, kc, ty, tw):
    bad, rel, _ = run(im, jm, km, kc, ty, tw=tw)
    assert bad == 0
    assert rel < 1e-12  # the stop-test sum: same terms, another order
