"""Pins of the oracle's pressure solve.

1. The reduced symmetric system.  PETSc is not available here, so the solve
   is checked against an independent statement of what it solves: the full
   DMDA system of compute_matrix / compute_rhs (pressure_solver.F90:234-441,
   ghost rows included, PETSc index order i_P = j_T - 1, j_P = i_T, k_P = k_T
   from pressure_solver.F90:195-204) is assembled here with scipy, its ghost
   unknowns are eliminated by a sparse Schur complement, the compatible part
   of the reduced right-hand side is solved directly, and the ghost values are
   recovered from the ghost rows.  The oracle's conjugate-gradient solution
   (plain and multigrid-preconditioned) must agree with it everywhere,
   ghost planes included, to the solver tolerance.
2. The multigrid preconditioner: the V-cycle is symmetric and positive
   definite on mean-free vectors (CG needs both), it contracts the residual
   of the operator restated here in numpy at a mesh-independent rate, and
   the preconditioned CG iteration count does not grow with the grid.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import binding as B


# ---------------------------------------------------------------------------
# 1. the full PETSc system, restated
# ---------------------------------------------------------------------------
def petsc_system(rhs_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt):
    """Matrix and right-hand side of compute_matrix / compute_rhs on the
    single-rank domain.  PETSc grid: i in 0..gdx-1 (periodic, TPLS j-1),
    j in 0..gdy+1 (TPLS i), k in 0..gdz+1; gdx = maxm-1, gdy = maxl-1,
    gdz = maxn-1.  Returns (A, b, index) with index[(i_T, j_T, k)] -> row."""
    gdx, gdy, gdz = maxm - 1, maxl - 1, maxn - 1
    nI, nJ, nK = gdx, gdy + 2, gdz + 2
    row = lambda i, j, k: i + nI * (j + nJ * k)
    rows, cols, vals = [], [], []
    b = np.zeros(nI * nJ * nK)

    def put(r, entries):
        for (i, j, k), v in entries:
            rows.append(r)
            cols.append(row(i % nI, j, k))
            vals.append(v)

    for k in range(nK):
        for j in range(nJ):
            for i in range(nI):
                r = row(i, j, k)
                # compute_rhs: b(i,j,k) = -RHS_p(j, i+1, k); inlet flux on j = 0
                b[r] = -rhs_p[j, i + 1, k]
                if j == 0 and 1 <= k <= gdz:
                    b[r] = (dx / dt) * (u_inlet[k - 1] - u3[0, i, k - 1]) / dy
                if (j == 0 or j == gdy + 1) and k == 0:
                    put(r, [((i, j, k), 1 / dz), ((i, j, k + 1), -1 / dz)])
                elif (j == 0 or j == gdy + 1) and k == gdz + 1:
                    put(r, [((i, j, k), -1 / dz), ((i, j, k - 1), 1 / dz)])
                elif j == 0:
                    put(r, [((i, j, k), 1 / dy), ((i, j + 1, k), -1 / dy)])
                elif j == gdy + 1:
                    put(r, [((i, j, k), -1 / dy), ((i, j - 1, k), 1 / dy)])
                elif k == 0:
                    put(r, [((i, j, k), 1 / dz), ((i, j, k + 1), -1 / dz)])
                elif k == gdz + 1:
                    put(r, [((i, j, k), -1 / dz), ((i, j, k - 1), 1 / dz)])
                else:
                    put(r, [((i, j, k - 1), -1 / dz**2), ((i, j - 1, k), -1 / dy**2),
                            ((i - 1, j, k), -1 / dx**2),
                            ((i, j, k), 2 * (1 / dx**2 + 1 / dy**2 + 1 / dz**2)),
                            ((i + 1, j, k), -1 / dx**2), ((i, j + 1, k), -1 / dy**2),
                            ((i, j, k + 1), -1 / dz**2)])
    n = nI * nJ * nK
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    index = {}
    for k in range(nK):
        for j in range(nJ):
            for i in range(nI):
                index[(j, i + 1, k)] = row(i, j, k)   # TPLS (i, j, k)
    interior = np.array([row(i, j, k) for k in range(1, gdz + 1) for j in range(1, gdy + 1)
                         for i in range(nI)])
    return A, b, index, interior


def direct_solution(rhs_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt, x_mean):
    A, b, index, I = petsc_system(rhs_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt)
    n = A.shape[0]
    G = np.setdiff1d(np.arange(n), I)
    AII, AIG = A[I][:, I], A[I][:, G]
    AGI, AGG = A[G][:, I], A[G][:, G].tocsc()
    # Schur complement: x_G = AGG^-1 (b_G - AGI x_I)
    S = (AII - AIG @ spla.spsolve(AGG, AGI.tocsc())).tocsr()
    g = b[I] - AIG @ spla.spsolve(AGG, b[G])
    assert np.allclose(S.toarray(), S.toarray().T, rtol=0, atol=1e-9 * abs(S).max())
    g = g - g.mean()    # the compatible part (constant null space)
    m = len(I)
    K = sp.bmat([[S, np.ones((m, 1))], [np.ones((1, m)), None]]).tocsc()
    sol = spla.spsolve(K, np.concatenate([g, [x_mean * m]]))
    x = np.zeros(n)
    x[I] = sol[:m]
    x[G] = spla.spsolve(AGG, b[G] - AGI @ x[I])
    pres = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    for (iT, jT, k), r in index.items():
        pres[iT, jT, k] = x[r]
    pres[:, 0, :] = pres[:, maxm - 1, :]     # periodic halo rows
    pres[:, maxm, :] = pres[:, 1, :]
    return pres


@pytest.mark.parametrize("pc", [0, 1, 2])
def test_solve_matches_full_petsc_system(pc):
    maxl, maxm, maxn = 13, 9, 11
    dx, dy, dz, dt = 1 / 10, 1 / 10, 1 / 10, 1e-3
    rng = np.random.default_rng(5)
    rhs_p = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    rhs_p[1:maxl, 1:maxm, 1:maxn] = rng.standard_normal((maxl - 1, maxm - 1, maxn - 1))
    u3 = np.asfortranarray(0.01 * rng.standard_normal((maxl + 1, maxm + 1, maxn - 1)))
    inl = np.ascontiguousarray(0.5 + 0.01 * rng.standard_normal(maxn - 1))
    pres = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    pres[1:maxl, 1:maxm, 1:maxn] = 0.1 * rng.standard_normal((maxl - 1, maxm - 1, maxn - 1))
    x_mean = pres[1:maxl, 1:maxm, 1:maxn].mean()
    reason, its, rnorm = B.pressure_solve(pres, rhs_p, u3, inl, maxl, maxm, maxn, dx, dy, dz, dt,
                                          rtol=1e-13, pc=pc)
    assert reason == 2
    want = direct_solution(rhs_p, u3, inl, maxl, maxm, maxn, dx, dy, dz, dt, x_mean)
    err = np.abs(pres - want).max() / np.abs(want).max()
    assert err < 1e-10, (pc, its, err)


@pytest.mark.parametrize("pc", [0, 2])
@pytest.mark.parametrize("where", ["rhs", "guess"])
def test_nan_input_is_reported_as_diverged_nan(pc, where):
    """A NaN in the right-hand side or in the initial guess reaches the norm
    of the initial residual, so the solve stops at once with PETSc's
    KSP_DIVERGED_NAN (-9), the code output_converged_reason prints as such
    (pressure_solver.F90:103-104), instead of reporting convergence."""
    maxl, maxm, maxn = 13, 9, 11
    rng = np.random.default_rng(6)
    rhs_p = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    rhs_p[1:maxl, 1:maxm, 1:maxn] = rng.standard_normal((maxl - 1, maxm - 1, maxn - 1))
    u3 = np.zeros((maxl + 1, maxm + 1, maxn - 1), order="F")
    inl = np.zeros(maxn - 1)
    pres = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    (rhs_p if where == "rhs" else pres)[4, 3, 5] = np.nan
    reason, its, _ = B.pressure_solve(pres, rhs_p, u3, inl, maxl, maxm, maxn, 0.1, 0.1, 0.1, 1e-3, pc=pc)
    assert reason == -9 and its == 0


def test_solve_detects_sign_and_axis_mistakes():
    """The direct solution is sensitive to the mistakes the elimination could
    make: a swapped x/y coefficient or a dropped inlet term moves it far
    outside the solver tolerance."""
    maxl, maxm, maxn = 13, 9, 11
    dx, dy, dz, dt = 1 / 10, 1 / 8, 1 / 12, 1e-3
    rng = np.random.default_rng(6)
    rhs_p = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    rhs_p[1:maxl, 1:maxm, 1:maxn] = rng.standard_normal((maxl - 1, maxm - 1, maxn - 1))
    u3 = np.asfortranarray(0.01 * rng.standard_normal((maxl + 1, maxm + 1, maxn - 1)))
    inl = np.ascontiguousarray(0.5 + 0.5 * rng.standard_normal(maxn - 1))
    pres = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    reason, _, _ = B.pressure_solve(pres, rhs_p, u3, inl, maxl, maxm, maxn, dx, dy, dz, dt,
                                    rtol=1e-13)
    assert reason == 2
    want = direct_solution(rhs_p, u3, inl, maxl, maxm, maxn, dx, dy, dz, dt, 0.0)
    assert np.abs(pres - want).max() / np.abs(want).max() < 1e-10
    swapped = direct_solution(rhs_p, u3, inl, maxl, maxm, maxn, dy, dx, dz, dt, 0.0)
    no_inlet = direct_solution(rhs_p, u3, 0 * inl, maxl, maxm, maxn, dx, dy, dz, dt, 0.0)
    for other in (swapped, no_inlet):
        assert np.abs(pres - other).max() / np.abs(want).max() > 1e-3


# ---------------------------------------------------------------------------
# 2. the multigrid preconditioner
# ---------------------------------------------------------------------------
def apply_A(x, cx, cy, cz, diag):
    """Reduced operator on x[k, j, i]: Neumann (reflecting) in i and k,
    periodic in j."""
    xm = np.concatenate([x[:, :, :1], x[:, :, :-1]], axis=2)
    xp = np.concatenate([x[:, :, 1:], x[:, :, -1:]], axis=2)
    zm = np.concatenate([x[:1], x[:-1]], axis=0)
    zp = np.concatenate([x[1:], x[-1:]], axis=0)
    ym, yp = np.roll(x, 1, axis=1), np.roll(x, -1, axis=1)
    return diag * x - cx * (xm + xp) - cy * (ym + yp) - cz * (zm + zp)


GRIDS = [(16, 8, 12, (1.0, 1.0, 1.0)), (32, 16, 32, (1.0, 4.0, 0.25)), (36, 16, 28, (2.0, 1.0, 3.0))]


@pytest.mark.parametrize("nu", [1, 2])
@pytest.mark.parametrize("grid", GRIDS)
def test_vcycle_symmetric_positive_definite(grid, nu):
    nx, ny, nz, (cx, cy, cz) = grid
    diag = 2 * (cx + cy + cz)
    rng = np.random.default_rng(nx + nu)
    for _ in range(3):
        u = rng.standard_normal(nx * ny * nz)
        v = rng.standard_normal(nx * ny * nz)
        u -= u.mean()
        v -= v.mean()
        Mu, _ = B.mg_apply(u, nx, ny, nz, cx, cy, cz, diag, nu)
        Mv, _ = B.mg_apply(v, nx, ny, nz, cx, cy, cz, diag, nu)
        assert abs(u @ Mv - v @ Mu) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Mv)
        assert u @ Mu > 0 and v @ Mv > 0


def test_levels():
    # coarsening stops when nx or nz turns odd, ny stops being a multiple of 4
    # or an extent would fall below 4
    one = lambda nx, ny, nz: B.mg_apply(np.zeros(nx * ny * nz), nx, ny, nz, 1, 1, 1, 6)[1]
    assert one(64, 16, 32) == 4          # 64x16x32 .. 8x2x4
    assert one(36, 16, 28) == 3          # .. 9x4x7
    assert one(512, 8, 8) == 3           # 256x4x4, 128x2x2
    assert one(512, 8, 6) == 2           # 256x4x3 (nz odd)
    assert one(9, 4, 7) == 1


@pytest.mark.parametrize("nu,bound", [(1, 0.5), (2, 0.3)])
def test_vcycle_contracts_at_mesh_independent_rate(nu, bound):
    """Stationary iteration x <- x + M (b - A x) with the mean removed; the
    residual of the numpy operator above shrinks by at most `bound` per cycle
    on two grid sizes, and the rate differs little between them."""
    rates = []
    for (nx, ny, nz) in [(32, 16, 32), (64, 32, 64)]:
        rng = np.random.default_rng(1)
        b = rng.standard_normal((nz, ny, nx))
        b -= b.mean()
        x = np.zeros_like(b)
        rn = [np.linalg.norm(b)]
        for _ in range(8):
            r = b - apply_A(x, 1.0, 1.0, 1.0, 6.0)
            z, _ = B.mg_apply(r.ravel(), nx, ny, nz, 1.0, 1.0, 1.0, 6.0, nu)
            z = z.reshape(b.shape)
            x += z - z.mean()
            rn.append(np.linalg.norm(b - apply_A(x, 1.0, 1.0, 1.0, 6.0)))
        rate = (rn[-1] / rn[2]) ** (1 / 6)
        assert rate < bound, (nx, rate)
        rates.append(rate)
    assert abs(rates[0] - rates[1]) < 0.05


@pytest.mark.parametrize("grid", [(9, 4, 7, (1.0, 1.0, 1.0)), (9, 6, 7, (2.0, 1.0, 3.0)), (7, 5, 9, (1.0, 4.0, 0.25)),
                                  (24, 6, 12, (1.0, 1.0, 1.0)), (2, 2, 3, (1.0, 1.0, 1.0))])
def test_coarsest_solve_is_the_exact_mean_free_solve(grid):
    """Where the grid does not coarsen, one V-cycle is the coarsest-level
    solve: by fast diagonalisation it is the exact solve on the mean-free
    subspace, A z = b - mean(b) with mean(z) = 0, checked against the numpy
    operator (even and odd periodic extents, anisotropic coefficients, the
    24x6x12 coarsest level of a 192x48x96 grid, extents of 2)."""
    nx, ny, nz, (cx, cy, cz) = grid
    diag = 2 * (cx + cy + cz)
    rng = np.random.default_rng(nx * ny + nz)
    b = rng.standard_normal((nz, ny, nx))
    z, nl = B.mg_apply(b.ravel(), nx, ny, nz, cx, cy, cz, diag)
    assert nl == 1
    z = z.reshape(b.shape)
    assert abs(z.mean()) < 1e-14 * np.abs(z).max()
    r = (b - b.mean()) - apply_A(z, cx, cy, cz, diag)
    assert np.linalg.norm(r) < 1e-12 * np.linalg.norm(b)


def _channel_solve(nx, ny, nz, pc, seed=3):
    maxl, maxm, maxn = nx + 1, ny + 1, nz + 1
    d = 1.0 / nz
    rng = np.random.default_rng(seed)
    rhs_p = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    rhs_p[1:maxl, 1:maxm, 1:maxn] = rng.standard_normal((nx, ny, nz))
    u3 = np.zeros((maxl + 1, maxm + 1, maxn - 1), order="F")
    inl = np.zeros(maxn - 1)
    pres = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    reason, its, _ = B.pressure_solve(pres, rhs_p, u3, inl, maxl, maxm, maxn, d, d, d, 1e-4,
                                      rtol=1e-10, pc=pc)
    assert reason == 2
    return pres, its


def test_mg_pcg_matches_plain_cg_and_is_mesh_independent():
    p0, its0 = _channel_solve(64, 16, 32, 0)
    its = {}
    for nu in (1, 2):
        p, its[nu] = _channel_solve(64, 16, 32, nu)
        assert np.abs(p - p0).max() < 1e-8 * np.abs(p0).max()
        assert its[nu] < its0 / 5
    _, its_big = _channel_solve(128, 32, 64, 2)
    assert its_big <= its[2] + 2
