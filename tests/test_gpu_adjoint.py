"""GPU adjoint: d lambda / d eps through the C-ABI (fame_eig_grad) and the
torch.autograd.Function built on it.

* Against dense FP64 eigenvectors (scipy.linalg.eigh of C^H C E = lambda B E
  with the restated sparse curl) on a tiny grid: Hellmann-Feynman with
  B-normalised vectors gives d lambda_j / d eps_e = -lambda_j |E_ej|^2.
* Against central finite differences of the eigensolver itself, for a smooth
  continuous permittivity eps = 1 + 11.96 sigmoid(4 z) on the cells, mapped to
  the Yee edges by averaging the two cells that share each edge; objective
  lambda_6 - lambda_5 summed over the wave vectors (bands counted from 1).
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import fame_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _smooth_field(shape, seed, sigma=2.0):
    """periodic Gaussian-smoothed N(0,1) field (z, y, x)"""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    kz, ky, kx = np.meshgrid(*k, indexing="ij")
    g = np.exp(-2 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    f = np.real(np.fft.ifftn(np.fft.fftn(z) * g))
    return f / f.std()


def test_grad_matches_dense_eigenvectors(gpu):
    import torch
    lo = F.build_lattice("simple_cubic", 6)
    fl = F.to_fame_lattice(lo)
    ol = O.lattice_from(fl)
    N = 6 ** 3
    f = _smooth_field((6, 6, 6), 11)
    eps = F.cell_to_yee(torch.tensor(1.0 + 11.96 / (1.0 + np.exp(-4.0 * f)))).numpy()
    k = np.array([0.21, -0.13, 0.07])
    nwant = 6
    r, c, v = O.curl_coo(ol, k)
    Cm = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).toarray()
    w, V = sl.eigh(Cm.conj().T @ Cm, np.diag(eps))
    first = int(np.argmax(w > 1e-8 * w.max()))
    lam = w[first:first + nwant]
    Vn = V[:, first:first + nwant]                     # B-normalised: V^H B V = I
    assert np.min(np.diff(lam)) > 1e-6 * lam[-1]       # no degeneracy at this k
    wts = np.random.default_rng(3).standard_normal(nwant)
    g_ref = -(np.abs(Vn) ** 2 * (lam * wts)[None, :]).sum(axis=1)

    S = F.Solver(fl, max_batch=1)
    S.set_eps(eps)
    S.set_wave_vectors(k)
    freq, _, E, prof = S.eigs(nwant, 18, want_fields=True)
    assert np.max(np.abs(freq[0] ** 2 - lam) / lam) < 1e-8
    g = S.eig_vjp(freq, E, wts[None, :]).cpu().numpy()
    assert np.abs(g - g_ref).max() < 1e-8 * np.abs(g_ref).max()
    S.close()


def _objective_setup(nx=12):
    import torch
    lo = F.build_lattice("simple_cubic", nx)
    fl = F.to_fame_lattice(lo)
    z0 = torch.tensor(_smooth_field((nx, nx, nx), 4), dtype=torch.float64)
    ks = np.array([[0.25, 0.125, 0.0], [0.1, 0.3, 0.05]])
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(np.ones(3 * nx ** 3))
    S.set_wave_vectors(ks)

    def J(z):
        eps = F.cell_to_yee(1.0 + 11.96 * torch.sigmoid(4.0 * z))
        lam = F.band_eigenvalues(S, eps, 8, 24)
        return (lam[:, 5] - lam[:, 4]).sum()
    return S, z0, J


def test_autograd_directional_fd(gpu):
    import torch
    S, z0, J = _objective_setup()
    z = z0.clone().requires_grad_(True)
    val = J(z)
    val.backward()
    g = z.grad.clone()
    d = torch.tensor(np.random.default_rng(7).standard_normal(z0.shape))
    t = 1e-4
    with torch.no_grad():
        fd = (J(z0 + t * d) - J(z0 - t * d)).item() / (2 * t)
    an = float((g * d).sum())
    assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
    S.close()


def test_autograd_per_cell_fd(gpu):
    import torch
    S, z0, J = _objective_setup()
    z = z0.clone().requires_grad_(True)
    J(z).backward()
    g = z.grad.clone()
    rng = np.random.default_rng(8)
    flat = np.argsort(-np.abs(g.numpy()).ravel())[:40]
    cells = rng.choice(flat, 4, replace=False)          # cells with a sizeable derivative
    h = 1e-3
    for c in cells:
        e = torch.zeros(z0.numel(), dtype=torch.float64)
        e[c] = 1.0
        e = e.reshape(z0.shape)
        with torch.no_grad():
            fd = (J(z0 + h * e) - J(z0 - h * e)).item() / (2 * h)
        an = float(g.reshape(-1)[c])
        assert abs(fd - an) <= 1e-3 * abs(an) + 1e-9, (int(c), fd, an)
    S.close()


def test_restore_matches_eigs_fields(gpu):
    """fame_restore from the reduced Ritz vectors reproduces fame_eigs' fields,
    including the Gamma rule (zero modes prepended)."""
    from cases import KPTS, make
    lo, fl, ol, eps = make("fcc")
    ks = np.stack([KPTS["k"], KPTS["gamma"]])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    freq, evr, E, _ = S.eigs(6, 20, want_reduced=True, want_fields=True)
    E2 = S.restore(evr)
    assert float((E2 - E).abs().max()) <= 1e-13 * float(E.abs().max())
    S.close()


def test_grad_of_degenerate_cluster_is_the_trace_derivative(gpu):
    """A structure with cubic symmetry (SC, sphere at the cell centre, uniform
    eps) at k along z: bands come in exactly degenerate pairs.  Inside a
    cluster only the trace is differentiable; fame_eig_grad replaces the
    weights of a cluster by their mean, so that
      * weights (1, 1) on a pair give the derivative of the pair's trace, which
        any orthonormal basis of the dense eigenspace reproduces, and
      * weights (1, 0) give half of it."""
    lo = F.build_lattice("simple_cubic", 6)
    fl = F.to_fame_lattice(lo)
    ol = O.lattice_from(fl)
    N = 6 ** 3
    mat = F.Material(sphere_centers=[[(0.5, 0.5, 0.5)]], sphere_radius=[[0.3]])
    eps = F.material_eps(lo, mat, [13.0], 1.0)
    k = np.array([0.0, 0.0, 0.17])
    nwant = 6
    r, c, v = O.curl_coo(ol, k)
    Cm = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).toarray()
    w, V = sl.eigh(Cm.conj().T @ Cm, np.diag(eps))
    first = int(np.argmax(w > 1e-8 * w.max()))
    lam = w[first:first + nwant + 1]
    Vn = V[:, first:first + nwant + 1]
    # the clusters among the wanted bands, from the dense spectrum
    clusters, i = [], 0
    while i < nwant:
        j = i + 1
        while j < nwant + 1 and abs(lam[j] - lam[i]) <= 1e-9 * lam[i]:
            j += 1
        clusters.append((i, j))
        i = j
    pairs = [(a, b) for a, b in clusters if b - a == 2 and b <= nwant]
    assert pairs, (lam, clusters)                         # the scenario this test is about
    a, b = pairs[0]

    S = F.Solver(fl, max_batch=1)
    S.set_eps(eps)
    S.set_wave_vectors(k)
    freq, _, E, _ = S.eigs(nwant, 20, want_fields=True)
    assert np.max(np.abs(freq[0] ** 2 - lam[:nwant]) / lam[:nwant]) < 1e-8
    trace_ref = -(np.abs(Vn[:, a:b]) ** 2).sum(axis=1) * lam[a]     # sum_j -lambda |E_j|^2 over the pair
    w11 = np.zeros((1, nwant)); w11[0, a:b] = 1.0
    w10 = np.zeros((1, nwant)); w10[0, a] = 1.0
    g11 = S.eig_vjp(freq, E, w11).cpu().numpy()
    g10 = S.eig_vjp(freq, E, w10).cpu().numpy()
    assert np.abs(g11 - trace_ref).max() < 1e-8 * np.abs(trace_ref).max()
    assert np.abs(g10 - 0.5 * g11).max() < 1e-12 * np.abs(g11).max()
    S.close()
