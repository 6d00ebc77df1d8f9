"""The oracle's supercell shift against an independent statement of the
canonical order, and against what any correct shift must satisfy: every
particle kept exactly once, each leaver in the supercell its multiMask code
points to (periodic), stayers in their own supercell, no holes below count.

Canonical order (shared by the HIP path): per supercell the leavers are
grouped by exchange code, ascending in slot inside a code; the supercell is
compacted on its own (with nn = count - L, the r-th leaver slot below nn
receives the r-th stayer at or above nn); then every supercell appends, for
codes 1..26, the leavers its neighbour sends towards it with that code."""
import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import setups

NSC = (2, 3, 2)  # supercells per axis


def code_dir(code):
    e = [code % 3, (code // 3) % 3, code // 9]
    return [-1 if v == 2 else v for v in e]


def random_frames(st, rng, leave_frac):
    nsc = int(np.prod(NSC))
    cap = st.species[0].frames() * 256
    f = {"counts": np.zeros(nsc, np.uint32), "pos": np.zeros((3, nsc * cap), np.float32),
         "mom": np.zeros((3, nsc * cap), np.float32), "w": np.zeros(nsc * cap, np.float32),
         "cell": np.zeros(nsc * cap, np.uint16), "mask": np.zeros(nsc * cap, np.uint8)}
    for sc in range(nsc):
        n = int(rng.integers(0, cap // 2))
        f["counts"][sc] = n
        s = slice(sc * cap, sc * cap + n)
        f["pos"][:, s] = rng.random((3, n), np.float32)
        f["mom"][:, s] = rng.standard_normal((3, n)).astype(np.float32)
        f["w"][s] = rng.random(n, np.float32) + 1.0
        f["cell"][s] = rng.integers(0, 256, n)
        m = np.ones(n, np.uint8)
        leave = rng.random(n) < leave_frac
        m[leave] = rng.integers(2, 28, int(leave.sum()))
        f["mask"][s] = m
    return f, cap


def brute_force(f, cap):
    """Slot lists of particle identities (slot order of set_frames) after the shift."""
    nsc = len(f["counts"])
    ids, masks = [], []
    nxt = 0
    for sc in range(nsc):
        n = int(f["counts"][sc])
        ids.append(list(range(nxt, nxt + n)))
        masks.append([int(v) for v in f["mask"][sc * cap:sc * cap + n]])
        nxt += n
    out = {}
    for sc in range(nsc):
        lv = [i for i, m in enumerate(masks[sc]) if m >= 2]
        for code in range(1, 27):
            out[(sc, code)] = [ids[sc][i] for i in lv if masks[sc][i] - 1 == code]
        nn = len(ids[sc]) - len(lv)
        srcs = [i for i in range(nn, len(ids[sc])) if masks[sc][i] == 1]
        arr = list(ids[sc])
        for r, i in enumerate(srcs):
            arr[lv[r]] = ids[sc][i]
        ids[sc] = arr[:nn]
    nx, ny, nz = NSC
    final = []
    for d in range(nsc):
        dx, dy, dz = d % nx, (d // nx) % ny, d // (nx * ny)
        lst = list(ids[d])
        for code in range(1, 27):
            t = code_dir(code)
            src = ((dx - t[0]) % nx) + nx * (((dy - t[1]) % ny) + ny * ((dz - t[2]) % nz))
            lst += out[(src, code)]
        final.append(lst)
    return final


@pytest.mark.parametrize("leave_frac", [0.0, 0.05, 0.5, 1.0])
def test_oracle_shift_matches_canonical_order(leave_frac):
    st = setups.thermal(grid=(8 * NSC[0], 8 * NSC[1], 4 * NSC[2]), ppc=1)
    rng = np.random.default_rng(int(leave_frac * 100) + 3)
    f, cap = random_frames(st, rng, leave_frac)
    ora = O.Oracle(st, 32, 1)
    ora.set_frames(0, f)
    before = {}
    nsc = len(f["counts"])
    nxt = 0
    for sc in range(nsc):
        for i in range(int(f["counts"][sc])):
            s = sc * cap + i
            before[nxt] = (sc, int(f["mask"][s]), f["pos"][:, s].copy(), f["mom"][:, s].copy(), f["w"][s],
                           int(f["cell"][s]))
            nxt += 1
    ora.shift(0)
    g = ora.get_frames(0)
    ids = ora.get_ids(0)
    want = brute_force(f, cap)
    nx, ny, _ = NSC
    seen = []
    for d in range(nsc):
        n = int(g["counts"][d])
        got = [int(v) for v in ids[d * cap:d * cap + n]]
        assert got == want[d], f"supercell {d}"
        assert np.all(g["mask"][d * cap:d * cap + n] == 1)
        assert np.all(g["mask"][d * cap + n:(d + 1) * cap] == 0)
        for i, pid in enumerate(got):
            sc0, m, pos, mom, w, cell = before[pid]
            s = d * cap + i
            if m >= 2:  # lands in the supercell its code points to
                t = code_dir(m - 1)
                x0, y0, z0 = sc0 % nx, (sc0 // nx) % ny, sc0 // (nx * ny)
                dest = ((x0 + t[0]) % NSC[0]) + nx * (((y0 + t[1]) % NSC[1]) + ny * ((z0 + t[2]) % NSC[2]))
                assert dest == d
            else:
                assert sc0 == d
            np.testing.assert_array_equal(g["pos"][:, s], pos)
            np.testing.assert_array_equal(g["mom"][:, s], mom)
            assert g["w"][s] == w and int(g["cell"][s]) == cell
        seen += got
    assert sorted(seen) == list(range(nxt))
    ora.close()
