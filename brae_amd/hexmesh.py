"""Structured hex meshes in OpenFOAM polyMesh order, and their finite-volume geometry.

A box of nx*ny*nz hexahedra numbered i fastest (blockMesh's order), internal faces sorted by owner
and then by neighbour (the upper-triangular order lduAddressing requires), boundary faces grouped by
patch. Optional two-sided wall grading on y and z and seeded jitter of the interior vertices make
the graded, non-orthogonal channels the pressure benchmarks run on.

The geometry follows OpenFOAM's primitiveMesh algorithms -- face centres and areas from the
triangle fan about the point average, cell centres and volumes from the pyramid decomposition about
the face-centre average, weights, deltaCoeffs, nonOrthDeltaCoeffs and the non-orthogonal
correction vectors (the reference's fv_geometry.cu:13-158) -- vectorised over faces with numpy.
"""
import numpy as np

PATCHES_CHANNEL = ("inlet", "outlet", "lowerWall", "upperWall", "back", "front")
PATCHES_CAVITY = ("left", "right", "bottom", "movingWall", "back", "front")


def splitmix64_uniform(seed, n):
    """n doubles in [0,1) from SplitMix64 (Steele/Lea/Flood 2014) with the 53-bit mantissa map,
    reproducible across compilers and platforms."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def graded_coords(n, length, ratio=1.0):
    """n cells on [0, length]; ratio > 1 grades both ends towards the walls: each half expands
    geometrically from the wall to the middle with last/first cell = ratio."""
    if ratio == 1.0 or n < 4:
        return np.linspace(0.0, length, n + 1)
    m = n // 2
    q = ratio ** (1.0 / (m - 1)) if m > 1 else 1.0
    h = q ** np.arange(m)
    if n % 2:
        h = np.concatenate([h, [h[-1] * q], h[::-1]])
    else:
        h = np.concatenate([h, h[::-1]])
    x = np.concatenate([[0.0], np.cumsum(h)])
    return x * (length / x[-1])


def box_points(nx, ny, nz, length=(1.0, 1.0, 1.0), grading=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    xs = [graded_coords(n, L, g) for n, L, g in zip((nx, ny, nz), length, grading)]
    X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
    # point index = i + (nx+1)*(j + (ny+1)*k): blockMesh's point order
    pts = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(), Z.transpose(2, 1, 0).ravel()], 1)
    if jitter > 0.0:
        npx, npy, npz = nx + 1, ny + 1, nz + 1
        i = np.arange(npx); j = np.arange(npy); k = np.arange(npz)
        spacing = []
        for a, n in zip(xs, (nx, ny, nz)):
            d = np.diff(a)
            s = np.empty(n + 1)
            s[0] = d[0]; s[-1] = d[-1]
            s[1:-1] = np.minimum(d[:-1], d[1:])
            spacing.append(s)
        K, J, I = np.meshgrid(k, j, i, indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        u = splitmix64_uniform(seed, 3 * pts.shape[0]).reshape(-1, 3)
        amp = np.stack([spacing[0][I], spacing[1][J], spacing[2][K]], 1) * jitter
        pts = pts + np.where(interior[:, None], (2.0 * u - 1.0) * amp, 0.0)
    return np.ascontiguousarray(pts)


def box_topology(nx, ny, nz, patch_names=PATCHES_CHANNEL):
    """(faces [nFaces,4] point labels, owner [nFaces], neighbour [nInternal], patches) with
    patches = [(name, startFace, nFaces)] in the order x-min, x-max, y-min, y-max, z-min, z-max."""
    npx, npy = nx + 1, ny + 1

    def P(i, j, k):
        return i + npx * (j + npy * k)

    c = np.arange(nx * ny * nz, dtype=np.int64)
    ci = c % nx
    cj = (c // nx) % ny
    ck = c // (nx * ny)
    owners, neis, quads = [], [], []
    # internal faces of each owner in order of neighbour: +x (c+1), +y (c+nx), +z (c+nx*ny)
    hx, hy, hz = ci < nx - 1, cj < ny - 1, ck < nz - 1
    i1, j0, k0 = ci + 1, cj, ck
    fx = np.stack([P(i1, j0, k0), P(i1, j0 + 1, k0), P(i1, j0 + 1, k0 + 1), P(i1, j0, k0 + 1)], 1)
    fy = np.stack([P(ci, cj + 1, ck), P(ci, cj + 1, ck + 1), P(ci + 1, cj + 1, ck + 1), P(ci + 1, cj + 1, ck)], 1)
    fz = np.stack([P(ci, cj, ck + 1), P(ci + 1, cj, ck + 1), P(ci + 1, cj + 1, ck + 1), P(ci, cj + 1, ck + 1)], 1)
    # interleave per owner: a [nCells, 3] slot table, masked, then flattened in row order
    slot_quads = np.stack([fx, fy, fz], 1)                      # [nC, 3, 4]
    slot_nei = np.stack([c + 1, c + nx, c + nx * ny], 1)        # [nC, 3]
    slot_mask = np.stack([hx, hy, hz], 1)
    own_int = np.repeat(c[:, None], 3, 1)[slot_mask]
    nei_int = slot_nei[slot_mask]
    quad_int = slot_quads[slot_mask]
    owners.append(own_int); neis.append(nei_int); quads.append(quad_int)
    patches = []
    start = own_int.size
    # boundary patches, outward normals
    def bnd(mask, q):
        return c[mask], q[mask]
    xmin_c, xmin_q = bnd(ci == 0, np.stack([P(ci, cj, ck), P(ci, cj, ck + 1), P(ci, cj + 1, ck + 1),
                                            P(ci, cj + 1, ck)], 1))
    xmax_c, xmax_q = bnd(ci == nx - 1, fx)
    ymin_c, ymin_q = bnd(cj == 0, np.stack([P(ci, cj, ck), P(ci + 1, cj, ck), P(ci + 1, cj, ck + 1),
                                            P(ci, cj, ck + 1)], 1))
    ymax_c, ymax_q = bnd(cj == ny - 1, fy)
    zmin_c, zmin_q = bnd(ck == 0, np.stack([P(ci, cj, ck), P(ci, cj + 1, ck), P(ci + 1, cj + 1, ck),
                                            P(ci + 1, cj, ck)], 1))
    zmax_c, zmax_q = bnd(ck == nz - 1, fz)
    for name, (bc, bq) in zip(patch_names, ((xmin_c, xmin_q), (xmax_c, xmax_q), (ymin_c, ymin_q),
                                            (ymax_c, ymax_q), (zmin_c, zmin_q), (zmax_c, zmax_q))):
        owners.append(bc); quads.append(bq)
        patches.append((name, start, bc.size))
        start += bc.size
    owner = np.concatenate(owners).astype(np.int32)
    faces = np.concatenate(quads).astype(np.int32)
    neighbour = np.concatenate(neis).astype(np.int32)
    return faces, owner, neighbour, patches


def quad_geometry(points, faces, owner, neighbour, nCells):
    """OpenFOAM geometry of a mesh whose faces are all quads (fv_geometry.cu:13-158), per face and
    cell, in FP64. Returns a dict of numpy arrays."""
    nF = faces.shape[0]
    nI = neighbour.size
    P = [points[faces[:, k]] for k in range(4)]
    fc = (((P[0] + P[1]) + P[2]) + P[3]) / 4.0
    sumN = np.zeros((nF, 3)); sumA = np.zeros(nF); sumAc = np.zeros((nF, 3))
    for k in range(4):
        a, b = P[k], P[(k + 1) % 4]
        cc = (a + b) + fc
        u = b - a
        v = fc - a
        n = np.stack([u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1], u[:, 2] * v[:, 0] - u[:, 0] * v[:, 2],
                      u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]], 1)
        ar = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
        sumN += n
        sumA += ar
        sumAc += ar[:, None] * cc
    Cf = ((1.0 / 3.0) * sumAc) / sumA[:, None]
    Sf = 0.5 * sumN
    magSf = np.sqrt(Sf[:, 0] * Sf[:, 0] + Sf[:, 1] * Sf[:, 1] + Sf[:, 2] * Sf[:, 2])
    del P, sumN, sumAc
    # cell centre estimate: owned faces in order then neighbour faces in order (a per-cell sum whose
    # order is face order within each pass; np.add.at accumulates sequentially)
    cEst = np.zeros((nCells, 3)); nCF = np.zeros(nCells)
    np.add.at(cEst, owner, Cf); np.add.at(nCF, owner, 1.0)
    np.add.at(cEst, neighbour, Cf[:nI]); np.add.at(nCF, neighbour, 1.0)
    cEst /= nCF[:, None]
    C = np.zeros((nCells, 3)); V = np.zeros(nCells)
    d = Cf - cEst[owner]
    pv = Sf[:, 0] * d[:, 0] + Sf[:, 1] * d[:, 1] + Sf[:, 2] * d[:, 2]
    pc = 0.75 * Cf + 0.25 * cEst[owner]
    np.add.at(C, owner, pv[:, None] * pc); np.add.at(V, owner, pv)
    d = cEst[neighbour] - Cf[:nI]
    pv = Sf[:nI, 0] * d[:, 0] + Sf[:nI, 1] * d[:, 1] + Sf[:nI, 2] * d[:, 2]
    pc = 0.75 * Cf[:nI] + 0.25 * cEst[neighbour]
    np.add.at(C, neighbour, pv[:, None] * pc); np.add.at(V, neighbour, pv)
    C = np.where(np.abs(V)[:, None] > 1e-300, C / V[:, None], cEst)
    V = V * (1.0 / 3.0)
    del cEst, d, pv, pc
    o, n = owner[:nI], neighbour
    sf, cf = Sf[:nI], Cf[:nI]
    dOwn = cf - C[o]; dNei = C[n] - cf
    SfdOwn = np.abs(sf[:, 0] * dOwn[:, 0] + sf[:, 1] * dOwn[:, 1] + sf[:, 2] * dOwn[:, 2])
    SfdNei = np.abs(sf[:, 0] * dNei[:, 0] + sf[:, 1] * dNei[:, 1] + sf[:, 2] * dNei[:, 2])
    den = SfdOwn + SfdNei
    weights = np.where(np.abs(den) > 1e-150, SfdNei / np.where(den == 0, 1, den), 0.5)
    delta = C[n] - C[o]
    magDelta = np.sqrt(delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1] + delta[:, 2] * delta[:, 2])
    deltaCoeffs = 1.0 / magDelta
    unit = sf / magSf[:nI, None]
    ud = unit[:, 0] * delta[:, 0] + unit[:, 1] * delta[:, 1] + unit[:, 2] * delta[:, 2]
    nonOrth = 1.0 / np.fmax(ud, 0.05 * magDelta)
    corrVec = unit - delta * nonOrth[:, None]
    # boundary deltaCoeffs 1/|nHat.(Cf - C)| (fv_patch.cu:160-165)
    bs = Sf[nI:]; bm = magSf[nI:]
    nh = bs / bm[:, None]
    bd = Cf[nI:] - C[owner[nI:]]
    bDelta = 1.0 / np.abs(bd[:, 0] * nh[:, 0] + bd[:, 1] * nh[:, 1] + bd[:, 2] * nh[:, 2])
    return dict(Cf=Cf, Sf=Sf, magSf=magSf, C=C, V=V, weights=weights, deltaCoeffs=deltaCoeffs,
                nonOrthDeltaCoeffs=nonOrth, corrVec=corrVec, bDeltaCoeffs=bDelta)


def box_mesh(nx, ny, nz, length=(1.0, 1.0, 1.0), grading=(1.0, 1.0, 1.0), jitter=0.0, seed=0,
             patch_names=PATCHES_CHANNEL, geometry=True):
    pts = box_points(nx, ny, nz, length, grading, jitter, seed)
    faces, owner, neighbour, patches = box_topology(nx, ny, nz, patch_names)
    m = dict(points=pts, faces=faces, owner=owner, neighbour=neighbour, patches=patches, nCells=nx * ny * nz,
             nInternal=neighbour.size, shape=(nx, ny, nz))
    if geometry:
        m.update(quad_geometry(pts, faces, owner, neighbour, nx * ny * nz))
    return m


def patch_slice(mesh, name):
    """slice of the boundary-face range (0-based within the boundary) for a patch"""
    for n, s, k in mesh["patches"]:
        if n == name:
            return slice(s - mesh["nInternal"], s - mesh["nInternal"] + k)
    raise KeyError(name)
