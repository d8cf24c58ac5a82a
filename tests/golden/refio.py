"""Writers/readers for the reference solver's on-disk formats (test support).

Only used to build inputs for, and read outputs of, the reference binary when
golden fixtures are generated, and by tests that check the host-side readers.
All binary files are raw native-endian int32/float64 with no record markers,
i fastest (reference: src/io/plot3d.cpp:443-541 grid, :715-833 solution,
:881-989 function; BC file src/io/io.cpp:957-1369).
"""
import numpy as np


def write_grid(path, blocks):
    """blocks: list of (x, y, z) arrays shaped (nk, nj, ni)."""
    with open(path, "wb") as f:
        np.array([len(blocks)], dtype=np.int32).tofile(f)
        for x, _, _ in blocks:
            nk, nj, ni = x.shape
            np.array([ni, nj, nk], dtype=np.int32).tofile(f)
        for x, y, z in blocks:
            for c in (x, y, z):
                np.ascontiguousarray(c, dtype=np.float64).tofile(f)


def write_function(path, blocks):
    """blocks: list of arrays shaped (nvar, nk, nj, ni)."""
    with open(path, "wb") as f:
        np.array([len(blocks)], dtype=np.int32).tofile(f)
        for q in blocks:
            nv, nk, nj, ni = q.shape
            np.array([ni, nj, nk, nv], dtype=np.int32).tofile(f)
        for q in blocks:
            np.ascontiguousarray(q, dtype=np.float64).tofile(f)


def write_solution(path, blocks, step=0, time=0.0):
    """blocks: list of arrays shaped (nv, nk, nj, ni); tau = {0, 1, step, time}."""
    with open(path, "wb") as f:
        np.array([len(blocks)], dtype=np.int32).tofile(f)
        for q in blocks:
            _, nk, nj, ni = q.shape
            np.array([ni, nj, nk], dtype=np.int32).tofile(f)
        for q in blocks:
            np.array([0.0, 1.0, float(step), float(time)], dtype=np.float64).tofile(f)
            np.ascontiguousarray(q, dtype=np.float64).tofile(f)


def read_solution(path, nv=5):
    """Returns (list of (nv, nk, nj, ni) arrays, list of tau arrays)."""
    raw = open(path, "rb").read()
    nb = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    dims = np.frombuffer(raw, np.int32, 3 * nb, 4).reshape(nb, 3)
    off = 4 + 12 * nb
    out, taus = [], []
    for ni, nj, nk in dims:
        taus.append(np.frombuffer(raw, np.float64, 4, off).copy())
        off += 32
        n = int(ni) * int(nj) * int(nk)
        out.append(np.frombuffer(raw, np.float64, nv * n, off).reshape(nv, nk, nj, ni).copy())
        off += 8 * nv * n
    return out, taus


def read_function(path):
    raw = open(path, "rb").read()
    nb = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    dims = np.frombuffer(raw, np.int32, 4 * nb, 4).reshape(nb, 4)
    off = 4 + 16 * nb
    out = []
    for ni, nj, nk, nv in dims:
        n = int(ni) * int(nj) * int(nk) * int(nv)
        out.append(np.frombuffer(raw, np.float64, n, off).reshape(nv, nk, nj, ni).copy())
        off += 8 * n
    return out


def write_bc(path, lines):
    """lines: iterable of (iblock, type, side, is, ie, js, je, ks, ke) with the
    index tokens already formatted as strings (e.g. '-0', '19')."""
    with open(path, "w") as f:
        for ln in lines:
            f.write(" ".join(str(t) for t in ln) + "\n")


def write_deck(path, items):
    """items: list of (name, value-string) in order."""
    with open(path, "w") as f:
        for k, v in items:
            f.write(f"{k} = {v}\n")


def sponge_lines(iblock, width, sides=("+xi", "-xi", "+eta", "-eta")):
    """Freund ambient sponges of `width` points on the requested faces of a 2-D
    block, in the BC-file index convention ('-k' means n-1-k)."""
    w = width - 1
    spec = {
        "+xi": ("0", str(w), "0", "-0", "0", "0"),
        "-xi": (f"-{w}", "-0", "0", "-0", "0", "0"),
        "+eta": ("0", "-0", "0", str(w), "0", "0"),
        "-eta": ("0", "-0", f"-{w}", "-0", "0", "0"),
    }
    return [(iblock, "sponge_freund_ambient", s) + spec[s] for s in sides]


def write_overture(path, ndim, dims, recv, widths, holes):
    """Overture explicit-interpolation file as read by
    src/io/overture.cpp:473-680.
    dims: [(ni, nj, nk)] per block; recv: list of (rblock, (i, j[, k]),
    dblock, (i, j[, k]) lower-left donor corner, coef[w3][w2][w1] flat),
    grouped by receiver block; widths: [(w1, w2, w3)] per receiver block;
    holes: list of (block, (i, j, k))."""
    nb = len(dims)
    recv = sorted(recv, key=lambda r: r[0])
    with open(path, "wb") as f:
        def ints(v):
            np.asarray(v, dtype=np.int32).tofile(f)
        ints([ndim, nb])
        ints([d[0] * d[1] * d[2] for d in dims])
        ints([sum(1 for r in recv if r[0] == b) for b in range(nb)])
        ints([r[0] for r in recv])
        ints([c for r in recv for c in r[1][:ndim]])
        ints([r[2] for r in recv])
        ints([c for r in recv for c in r[3][:ndim]])
        for b in range(nb):
            ints(widths[b])
            for r in recv:
                if r[0] == b:
                    np.asarray(r[4], dtype=np.float64).tofile(f)
        ints([sum(1 for h in holes if h[0] == b) for b in range(nb)])
        for b in range(nb):
            for h in holes:
                if h[0] == b:
                    ints(list(h[1]) + [0] * (3 - len(h[1])))


def time_matches(t, tau, report):
    """The solver time t against a fixture: tau[3] bit for bit, or, when the
    reference wrote a PLOT3D function file (a LEE_SCALAR run: no tau, so
    tau[3] is NaN), against the last report line's time, which the reference
    prints with std::scientific (6 digits)."""
    if not np.isnan(tau[3]):
        return t == float(tau[3])
    last = report.strip().splitlines()[-1]
    return f"{t:.6e}" == last.split("time:")[1].split(",")[0].strip()
