"""Pair schedules of the volume kernels (scratch_wadge_amd/csrc/wadges.hip,
vol_body), restated on the host for every compiled shape and checked for the two
properties the kernels rely on:

* coverage: every unordered volume-volume pair (diagonal included) and every
  (row, face point) volume-face pair of an element is evaluated exactly once;
* fixed accumulation order: within one LDS-add instruction (one step, one row
  slot, one lane group) no two lanes of an element add into the same address,
  so the order of the additions is set by the program alone.

The flux-differencing sums these pairs stand for are those of
okl/Euler3D.okl:2123-2523 and okl/Euler2D.okl:1112-1350; the GPU parity tests
check the sums themselves.
"""

import itertools

import pytest

WAVE = 64
# (DIM, NQ, NFQ) of the compiled instances (scratch_wadge_amd/csrc/wadges.hip, make_instance)
SHAPES = [(2, 3, 2), (2, 6, 3), (2, 12, 4), (2, 16, 5), (2, 25, 6),
          (3, 4, 3), (3, 11, 6), (3, 23, 12), (3, 44, 16)]


def shape(dim, nq, nfq):
    nfaces = dim + 1
    nfn = nfq * nfaces
    r = 2 if (dim, nq, nfq) == (3, 23, 12) else 1
    lpe = -(-nq // r)
    tpw = WAVE // lpe
    colvf = r == 1 and tpw == 1 and nfn == WAVE and nq < WAVE
    chunked = nq <= nfn
    groups = -(-nq // nfq)
    if r > 1:
        parts = nfaces
    elif chunked and dim == 2 and groups == 2:
        parts = nfaces
    else:
        parts = 2 if (chunked and nfaces == 4 and groups <= 2) else 1
    split = chunked and groups > nfaces // parts
    return dict(dim=dim, nq=nq, nfq=nfq, nfaces=nfaces, nfn=nfn, r=r, lpe=lpe, colvf=colvf,
                chunked=chunked, parts=parts, split=split)


def vv_schedule(s):
    """Volume-volume block: row i visits j = (i + t) mod NQ, t = 0..NQ/2 (the
    middle step of an even NQ only for i < NQ/2).  Yields (step, row slot, i, j)."""
    nq = s["nq"]
    for t in range(nq // 2 + 1):
        for r in range(s["r"]):
            for l in range(s["lpe"]):
                i = l + r * s["lpe"]
                if i >= nq or (2 * t == nq and i >= nq // 2):
                    continue
                yield (t, r), i, (i + t) % nq


def vf_schedule(s):
    """Volume-face block.  Yields (instruction key, row i, face point fn, the
    LDS address the step adds into: ('F', fn) for the face accumulator or ('V', i)
    for the row accumulator)."""
    nq, nfq, nfn, nfaces = s["nq"], s["nfq"], s["nfn"], s["nfaces"]
    if s["colvf"]:
        # lane L owns face point L and walks the rows; lanes below NQ and from
        # NQ on add into the rows in separate instructions
        for t in range(nq):
            for lane in range(WAVE):
                i = (lane % nq + t) % nq
                yield (t, lane >= nq), i, lane, ("V", i)
    elif s["r"] > 1:
        # part h is face h; at step m lane l visits point (l + m) mod NFQ for
        # each of its rows; one face add per lane and step
        for h in range(nfaces):
            for m in range(nfq):
                for l in range(s["lpe"]):
                    fn = h * nfq + (l + m) % nfq
                    for r in range(s["r"]):
                        i = l + r * s["lpe"]
                        if i < nq:
                            yield (h, m), i, fn, ("F", fn) if r == 0 else None
    elif s["chunked"]:
        fh = nfaces // s["parts"]
        for h in range(s["parts"]):
            for u in range(fh):
                for m in range(nfq):
                    for i in range(nq):
                        g, o = divmod(i, nfq)
                        face = h * fh + (u + g) % fh
                        fn = face * nfq + (o + m) % nfq
                        # groups meeting on a face add in one instruction each
                        yield (h, u, m, g if s["split"] else 0), i, fn, ("F", fn)
    else:
        L = max(nq, nfn)
        for t in range(L):
            for i in range(nq):
                fn = (i + t) % L
                if fn < nfn:
                    yield t, i, fn, ("F", fn)


@pytest.mark.parametrize("dim,nq,nfq", SHAPES)
def test_volume_volume_pairs_once_and_partner_adds_distinct(dim, nq, nfq):
    s = shape(dim, nq, nfq)
    seen = {}
    per_instr = {}
    for key, i, j in vv_schedule(s):
        pair = (min(i, j), max(i, j))
        seen[pair] = seen.get(pair, 0) + 1
        if i != j:   # the partner's share goes to its LDS row
            per_instr.setdefault(key, []).append(j)
    assert set(seen) == set(itertools.combinations_with_replacement(range(nq), 2))
    assert all(c == 1 for c in seen.values())
    for key, targets in per_instr.items():
        assert len(targets) == len(set(targets)), (key, targets)


@pytest.mark.parametrize("dim,nq,nfq", SHAPES)
def test_volume_face_pairs_once_and_adds_distinct(dim, nq, nfq):
    s = shape(dim, nq, nfq)
    seen = {}
    per_instr = {}
    for key, i, fn, addr in vf_schedule(s):
        seen[(i, fn)] = seen.get((i, fn), 0) + 1
        if addr is not None:
            per_instr.setdefault(key, []).append(addr)
    assert set(seen) == set(itertools.product(range(nq), range(s["nfn"])))
    assert all(c == 1 for c in seen.values())
    for key, addrs in per_instr.items():
        assert len(addrs) == len(set(addrs)), (key, addrs)


def test_triangle_face_walk_splits_its_two_groups():
    """Triangles with two lane groups walk one face at a time: both groups meet
    on the face's points, so one instruction for both would add twice into an
    address; the kernel adds per group."""
    for nq, nfq in ((3, 2), (6, 3)):
        s = shape(2, nq, nfq)
        assert s["parts"] == 3 and s["split"]
        per_step = {}
        for key, i, fn, addr in vf_schedule(s):
            per_step.setdefault(key[:3], []).append(addr)
        assert any(len(a) != len(set(a)) for a in per_step.values())


def test_one_element_walk_needs_its_two_lane_groups():
    """At Nq = 44, Nf = 64 the face-point walk's lanes L and L + 44 meet in the
    same row at every step: one instruction for all 64 lanes would add twice into
    20 rows, which is why the kernel splits the additions by lane group."""
    s = shape(3, 44, 16)
    assert s["colvf"]
    for t in range(s["nq"]):
        rows = [(lane % 44 + t) % 44 for lane in range(WAVE)]
        assert len(rows) - len(set(rows)) == WAVE - 44
