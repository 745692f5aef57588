"""CPU restatement of ODTbrain's sinogram preparation (TEST INFRASTRUCTURE:
a checker, never imported by the product package).

sinogram_as_rytov / sinogram_as_radon / align_unwrapped / divmod_neg restate
odtbrain/_prepare_sino.py:7-182.  The 3D path calls
skimage.restoration.unwrap_phase (scikit-image, absent here); unwrap_2d
below restates the algorithm that function implements, Herraez, Burton,
Lalor and Gdeisat, "Fast two-dimensional phase-unwrapping algorithm based on
sorting by reliability following a noncontinuous path", Applied Optics 41
(2002) 7437: pixel reliability from wrapped second differences, edges
between 4-neighbours with reliability the sum of their pixels', edges
processed in ascending order, groups merged with the smaller one shifted by
whole turns.  Two choices are this restatement's own and fixed: border
pixels (whose second differences are undefined) get reliability 1e7 (the
published code adds a random number to it), and equal edge reliabilities are
taken in creation order (horizontal edges row by row, then vertical).  For
phase fields without residues every processing order gives the same
unwrapped field up to one whole-turn offset per image, which alignment then
removes; that is what the reference's fixtures (processing__test_sino_rytov,
processing__test_sino_radon) and the tests here pin.
"""
import numpy as np

TWOPI = 2 * np.pi
BORDER_RELIABILITY = 1e7


def divmod_neg(a, b):
    """divmod with the remainder closest to zero (_prepare_sino.py:49-57)."""
    q, r = divmod(a, b)
    sr = np.sign(r)
    if np.abs(r) > b / 2:
        q += sr
        r -= b * sr
    return q, r


def _mode_smallest(steps):
    """scipy.stats.mode along axis 0 (most frequent value, the smallest on
    ties), exact float comparisons."""
    out = np.empty(steps.shape[1])
    for j in range(steps.shape[1]):
        vals, counts = np.unique(steps[:, j], return_counts=True)
        out[j] = vals[np.argmax(counts)]
    return out


def align_unwrapped(sino):
    """In place (_prepare_sino.py:7-46)."""
    samples = []
    if sino.ndim == 2:
        for c in (0, 1, 2, -1, -2):
            samples.append(sino[:, c])
    else:
        for iy, ix in ((0, 0), (0, -1), (-1, 0), (-1, -1), (0, 1)):
            samples.append(sino[:, iy, ix])
    steps = np.zeros((len(samples), samples[0].shape[0]))
    for i, s in enumerate(samples):
        steps[i] = s - np.unwrap(s)
    remove = _mode_smallest(steps)
    minimum = divmod_neg(np.min(sino), TWOPI)[0]
    remove += minimum * TWOPI
    for i in range(len(sino)):
        sino[i] -= remove[i]


def _wrap(d):
    if d > np.pi:
        return d - TWOPI
    if d < -np.pi:
        return d + TWOPI
    return d


def _find_wrap(v1, v2):
    d = v1 - v2
    if d > np.pi:
        return -1
    if d < -np.pi:
        return 1
    return 0


def unwrap_2d(img):
    """Reliability-sorted unwrapping of one wrapped phase image (Herraez et
    al. 2002), pure Python: small images only."""
    img = np.asarray(img, dtype=float)
    H, W = img.shape
    rel = np.full((H, W), BORDER_RELIABILITY)
    for i in range(1, H - 1):
        for j in range(1, W - 1):
            c = img[i, j]
            h = _wrap(img[i, j - 1] - c) - _wrap(c - img[i, j + 1])
            v = _wrap(img[i - 1, j] - c) - _wrap(c - img[i + 1, j])
            d1 = _wrap(img[i - 1, j - 1] - c) - _wrap(c - img[i + 1, j + 1])
            d2 = _wrap(img[i - 1, j + 1] - c) - _wrap(c - img[i + 1, j - 1])
            rel[i, j] = h * h + v * v + d1 * d1 + d2 * d2
    edges = []
    for i in range(H):
        for j in range(W - 1):
            edges.append((rel[i, j] + rel[i, j + 1], i * W + j, i * W + j + 1))
    for i in range(H - 1):
        for j in range(W):
            edges.append((rel[i, j] + rel[i + 1, j], i * W + j, (i + 1) * W + j))
    order = sorted(range(len(edges)), key=lambda k: (edges[k][0], k))
    flat = img.ravel()
    inc = merge_groups(H * W, [(edges[k][1], edges[k][2],
                                _find_wrap(flat[edges[k][1]], flat[edges[k][2]]))
                               for k in order])
    return (flat + TWOPI * inc).reshape(H, W)


def merge_groups(n, edges):
    """The merge sequence of the published algorithm: `edges` in sorted
    order as (p1, p2, e), e the whole-turn step find_wrap(phase[p1],
    phase[p2]); returns every pixel's whole-turn increment."""
    inc = np.zeros(n, dtype=np.int64)
    group = list(range(n))             # group id (head pixel) of every pixel
    members = {p: [p] for p in range(n)}
    for p1, p2, e in edges:
        g1, g2 = group[p1], group[p2]
        if g1 == g2:
            continue
        if len(members[g2]) == 1:        # p2 alone joins p1's group
            inc[p2] = inc[p1] - e
            members[g1].append(p2)
            group[p2] = g1
            del members[g2]
        elif len(members[g1]) == 1:      # p1 alone joins p2's group
            inc[p1] = inc[p2] + e
            members[g2].append(p1)
            group[p1] = g2
            del members[g1]
        elif len(members[g1]) > len(members[g2]):
            d = inc[p1] - e - inc[p2]
            for q in members[g2]:
                inc[q] += d
                group[q] = g1
            members[g1].extend(members.pop(g2))
        else:
            d = inc[p2] + e - inc[p1]
            for q in members[g1]:
                inc[q] += d
                group[q] = g2
            members[g2].extend(members.pop(g1))
    return inc


def _unwrapped_phase(phase):
    if phase.ndim == 2:
        return np.unwrap(phase, axis=-1)
    out = np.empty_like(phase)
    for ii in range(len(phase)):
        out[ii] = unwrap_2d(phase[ii])
    return out


def sinogram_as_radon(uSin, align=True):
    """Unwrapped phase of a complex sinogram (_prepare_sino.py:60-102)."""
    phiR = _unwrapped_phase(np.angle(uSin))
    if align:
        align_unwrapped(phiR)
    return phiR


def sinogram_as_rytov(uSin, u0=1, align=True):
    """Rytov phase u0 * (log|uSin/u0| + 1j * unwrapped angle(uSin/u0))
    (_prepare_sino.py:105-182)."""
    phiR = np.angle(uSin / u0)
    lna = np.log(np.absolute(uSin / u0))
    phiR = _unwrapped_phase(phiR)
    if align:
        align_unwrapped(phiR)
    return u0 * (1j * phiR + lna)


def processing_test_sino(rytov=False):
    """The sinogram of tests/test_processing.py:24-32."""
    ln, a = 10, 2
    sino = np.arange(ln * ln * a).reshape(a, ln, ln) / (ln * ln * a) * \
        np.exp(1j * np.arange(ln * ln * a).reshape(a, ln, ln))
    if rytov:
        sino[0, 0, 0] = .1
    return sino
