"""Writes tests/golden/refgrad/ref_autograd.npz: values and autograd gradients of the reference path stage
(feffgpu's GPUPathStage on the CPU in complex128, src/feffgpu/stage.py:96-388) for path geometries that only the
reference's own conventions fix: exactly collinear legs (beta = 0 and pi), legs along the z axis (the regularised
trig at rxy < 1e-6 and arg's (0, 0) case, stage.py:147-168), FEFF's W path list, and the random 2-8-leg walks over a
bcc lattice of tests/test_pathfinder.py:193-213 (displaced and undisplaced).

Run in the development container only (it imports the reference from /root/reference/src; nothing of it is stored
here but these vectors):

    python tests/golden/make_ref_autograd.py

Stored per group g (all paths of a group have the same number of legs): coords_g (B, n, 3) A, ipots_g (B, n),
cchi_g (B, E) complex, reff_g (B,), chi_g (B, NK) on q = 0..18 step 0.05 ("feff" convention), and the gradients
with respect to coords of two losses: gc_g of sum_e Re(conj(wc_e) cchi_e) + wr reff, and gq_g of sum_q wq_q chi_q,
with the weights wc (E,), wr, wq (NK,) stored once.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
HERE = Path(__file__).resolve().parent
OUT = HERE / "refgrad" / "ref_autograd.npz"
PHASE = HERE / "W" / "phase.pad"


def read_paths_dat(path):
    """(nleg, coords (n, 3) with the absorber first, ipots) of every path of FEFF's paths.dat."""
    lines = Path(path).read_text().splitlines()
    out, i = [], 0
    while i < len(lines):
        if "index, nleg" not in lines[i]:
            i += 1
            continue
        nleg = int(lines[i].split()[1])
        rows = [lines[i + 2 + j].split() for j in range(nleg)]
        pts = np.array([[float(v) for v in r[:3]] for r in rows])
        ip = [int(r[3]) for r in rows]
        coords = np.concatenate([pts[-1:], pts[:-1]])          # absorber (listed last) first
        coords = coords - coords[0]
        out.append((nleg, coords, np.array([ip[-1]] + ip[:-1])))
        i += 2 + nleg
    return out


def lattice_walk(rng, nleg, a=3.1652, sd=0.06):
    steps = np.vstack([np.array([[i, j, l] for i in (-.5, .5) for j in (-.5, .5) for l in (-.5, .5)]),
                       np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])]) * a
    while True:
        site, sites = np.zeros(3), []
        for _ in range(nleg - 1):
            site = site + steps[rng.integers(len(steps))]
            sites.append(site.copy())
        if np.linalg.norm(sites[-1]) > 1e-6:
            break
    moved, shift = [np.zeros(3)], {}
    for s in sites:
        key = tuple(np.round(s, 3))
        if np.linalg.norm(s) > 1e-6 and key not in shift:
            shift[key] = rng.normal(0, sd, 3)
        moved.append(s + shift.get(key, 0.0))
    return np.array(moved), np.array([0] + [0 if np.linalg.norm(s) < 1e-6 else 1 for s in sites])


def singular_paths():
    """Exactly collinear and axis-aligned geometries (W bcc distances), absorber at the origin."""
    a = 3.1652
    d1, d2 = a * np.sqrt(3) / 2, a
    z = lambda t: np.array([0.0, 0.0, t])
    diag = lambda t: np.full(3, t / np.sqrt(3))
    P = [
        (np.array([z(0), z(d2)]), [0, 1]),                                      # 2 legs along z (rxy = 0)
        (np.array([z(0), diag(d1)]), [0, 1]),                                   # 2 legs along the body diagonal
        (np.array([z(0), z(d2), z(2 * d2)]), [0, 1, 1]),                         # 3 legs, focusing, beta = 0 and pi
        (np.array([z(0), diag(d1), diag(2 * d1)]), [0, 1, 1]),                   # the same on the diagonal
        (np.array([z(0), z(-d2), z(d2)]), [0, 1, 1]),                            # through the absorber's line
        (np.array([z(0), z(d2), z(0), z(d2)]), [0, 1, 0, 1]),                    # 4 legs, absorber in the middle
        (np.array([z(0), z(d2), z(2 * d2), z(d2)]), [0, 1, 1, 1]),               # 4 legs, there and back
        (np.array([z(0), diag(d1), diag(2 * d1), diag(d1)]), [0, 1, 1, 1]),
        (np.array([z(0), z(d2), np.array([d2, 0, d2]), np.array([d2, 0, 0])]), [0, 1, 1, 1]),   # right angles in xz
        (np.array([z(0), np.array([0, d2, 0]), np.array([0, d2, d2]), z(d2)]), [0, 1, 1, 1]),
        (np.array([z(0), z(d2), z(2 * d2), z(3 * d2), z(2 * d2)]), [0, 1, 1, 1, 1]),   # 5 legs on a line
        (np.array([z(0), z(d2), z(0), z(-d2), z(0), z(d2)]), [0, 1, 0, 1, 0, 1]),      # 6 legs through the absorber
    ]
    return [(len(c), c, np.array(ip)) for c, ip in P]


def main():
    sys.path.insert(0, str(REF))
    import torch
    from feffgpu.stage import GPUPathStage
    torch.set_default_dtype(torch.float64)
    st = GPUPathStage(str(PHASE), device="cpu", dtype=torch.complex128)
    q = np.round(np.arange(0.0, 18.0 + 1e-9, 0.05), 10)
    S = st.linear_matrix(q)
    rng = np.random.default_rng(2026)
    E = st.E
    wc = (rng.normal(size=E) + 1j * rng.normal(size=E)) * np.exp(-0.08 * np.arange(E))
    wr = float(rng.normal())
    wq = rng.normal(size=len(q)) * q ** 2
    groups = {}

    def add(name, paths):
        by = {}
        for n, c, ip in paths:
            by.setdefault(n, []).append((c, ip))
        for n, lst in sorted(by.items()):
            groups[f"{name}_n{n}"] = (np.stack([c for c, _ in lst]), np.stack([ip for _, ip in lst]))

    add("singular", singular_paths())
    add("W", read_paths_dat(HERE / "W" / "paths.dat"))
    wr_rng = np.random.default_rng(5)
    walks = []
    for n in range(2, 9):
        for sd in (0.06, 0.0):
            for _ in range(2):
                c, ip = lattice_walk(wr_rng, n, sd=sd)
                walks.append((n, c, ip))
    add("walk", walks)
    out = dict(q=q, wc=wc, wr=np.array(wr), wq=wq)
    # Gradients are taken through torch.utils.checkpoint (use_reentrant=False), as the reference's own callers do
    # (src/feffgpu/exafs/forward.py:158-159, 683): the stage's _clm assigns into slices of tensors it has already
    # used, which a plain autograd pass refuses (version counter); the recomputation under checkpoint gives the
    # gradient the reference's forward models use.
    from torch.utils.checkpoint import checkpoint
    wct, wqt = torch.tensor(wc), torch.tensor(wq)
    for g, (coords, ipots) in groups.items():
        ip = torch.tensor(ipots)
        with torch.no_grad():
            cc, reff = st.cchi(torch.tensor(coords), ip)
            chi = st.chi_fine(cc, reff, q, S, "feff")

        def loss_c(x):
            c1, r1 = st.cchi(x, ip)
            return (wct.conj()[None] * c1).real.sum() + wr * r1.sum()

        def loss_q(x):
            c1, r1 = st.cchi(x, ip)
            return (st.chi_fine(c1, r1, q, S, "feff") * wqt[None]).sum()
        x = torch.tensor(coords, requires_grad=True)
        (g1,) = torch.autograd.grad(checkpoint(loss_c, x, use_reentrant=False), x)
        x2 = torch.tensor(coords, requires_grad=True)
        (g2,) = torch.autograd.grad(checkpoint(loss_q, x2, use_reentrant=False), x2)
        out[f"coords_{g}"] = coords
        out[f"ipots_{g}"] = ipots
        out[f"cchi_{g}"] = cc.numpy()
        out[f"reff_{g}"] = reff.numpy()
        out[f"chi_{g}"] = chi.numpy()
        out[f"gc_{g}"] = g1.numpy()
        out[f"gq_{g}"] = g2.numpy()
        print(g, coords.shape, "finite:", bool(np.isfinite(g1.numpy()).all() and np.isfinite(g2.numpy()).all()), flush=True)
    out["groups"] = np.array(sorted(groups))
    OUT.parent.mkdir(exist_ok=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
