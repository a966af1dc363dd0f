"""CPU restatement of the affine-body (ABD) linear-system slice, in numpy.

TEST INFRASTRUCTURE ONLY: the checker for the library's ABD kernels; the
library never calls it.

Follows:
- ABDJacobi (StiffGIPC/abd_system/abd_jacobi_matrix.h:7-60): body vertex
  x_i = p + (a1.xbar_i, a2.xbar_i, a3.xbar_i), q = (p, a1, a2, a3);
- the shape energy sum_i (|a_i|^2 - 1)^2 + 2 sum_{i<j} (a_i.a_j)^2
  (abd_system/details/abd_energy.inl:9-27), i.e. |A A^T - I|_F^2;
- _cal_abd_body_gradient_and_hessian
  (abd_system/abd_system_function/setup_abd_system_gradient_and_hessian.cu:
  201-345): G = M (q - q~) + kappa V dt^2 dS, H = M + PSD(kappa V dt^2 d2S)
  on the 9x9 affine block; fixed bodies G = 0, H = M;
- write_abd_body_hessian (:132-141): 10 upper 3x3 blocks per body;
- write_barrier_hessian (:87-128) with write_triplet_cv2 (:40-75): a
  body-body vertex block (i, j) becomes J_i^T H J_j as 16 blocks, those
  below the diagonal dropped; the ABD-FEM coupling (:521-580): the 4 blocks
  of J_i^T H; FEM-FEM blocks shifted to row v - nav + 4 nb, lower ones zeroed
  (StiffGIPC/GIPC.cu:10180-10215); blocks touching fixed bodies zero.  The
  reference applies these to its FULL vertex-space storage (SymGH is off,
  linear_system/linear_system/global_matrix.h:6: every vertex pair of a
  contact appears as (i, j) and (j, i)); map_contacts_full_storage restates
  that, map_contacts maps the upper storage the library uses (a same-body
  pair i < j -> upper blocks of X + X^T);
- _cal_abd_system_barrier_gradient (:348-422): grad_q += J_i^T g_i, fixed
  bodies skipped;
- cal_abd_kinetic_energy / cal_abd_shape_energy (cal_abd_energy.cu:6-140);
- the 12x12 block-Jacobi (abd_preconditioner.cu:15-38, :612-655).

The checks that pin it (tests/test_oracle_abd.py): central differences of
the energy against the gradient and of the gradient against the Hessian,
closed forms (S = 0 with zero gradient on rotations, S(sI) = 3 (s^2 - 1)^2),
the mapped contact blocks against a dense J^T H J built from the assembled
global J, the upper-storage mapping against the reference's full-storage
procedure, and J^T g against the dense J.
"""
import numpy as np


def jacobian(xbar):
    """3x12 J of one body vertex with rest coordinates xbar."""
    J = np.zeros((3, 12))
    J[:, :3] = np.eye(3)
    for k in range(3):
        J[k, 3 + 3 * k:6 + 3 * k] = xbar
    return J


def positions(q, body_of, xbar):
    """(nav, 3) body vertex positions J_i q."""
    q = np.asarray(q, np.float64).reshape(-1, 12)
    Q = q[body_of]
    A = Q[:, 3:].reshape(-1, 3, 3)              # rows a1, a2, a3
    return Q[:, :3] + np.einsum("nkj,nj->nk", A, xbar)


def shape_energy(a):
    A = np.asarray(a, np.float64).reshape(3, 3)
    D = A @ A.T
    return float(sum((D[i, i] - 1.0) ** 2 for i in range(3)) + 2.0 * (D[0, 1] ** 2 + D[0, 2] ** 2 + D[1, 2] ** 2))


def shape_gradient(a):
    A = np.asarray(a, np.float64).reshape(3, 3)
    D = A @ A.T
    # d/da_i = 4 (a_i.a_i - 1) a_i + 4 sum_{j != i} (a_i.a_j) a_j
    G = 4.0 * (D - np.eye(3)) @ A
    return G.reshape(9)


def shape_hessian(a):
    A = np.asarray(a, np.float64).reshape(3, 3)
    D = A @ A.T
    H = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            if i == j:
                blk = 4.0 * (D[i, i] - 1.0) * np.eye(3) + 8.0 * np.outer(A[i], A[i])
                for m in range(3):
                    if m != i:
                        blk += 4.0 * np.outer(A[m], A[m])
            else:
                blk = 4.0 * D[i, j] * np.eye(3) + 4.0 * np.outer(A[j], A[i])
            H[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
    return H


def make_psd(H):
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    return (V * np.maximum(w, 0.0)) @ V.T


def body_system(q, qt, M, vol, fixed, kappa, dt):
    """Per body: G (nb, 12) and the full 12x12 H (nb, 12, 12)."""
    q = np.asarray(q, np.float64).reshape(-1, 12)
    qt = np.asarray(qt, np.float64).reshape(-1, 12)
    M = np.asarray(M, np.float64).reshape(-1, 12, 12)
    nb = q.shape[0]
    G = np.zeros((nb, 12)); H = np.zeros((nb, 12, 12))
    for b in range(nb):
        H[b] = M[b]
        if fixed is not None and fixed[b]:
            continue
        w = kappa * vol[b] * dt * dt
        G[b] = M[b] @ (q[b] - qt[b])
        G[b, 3:] += w * shape_gradient(q[b, 3:])
        H[b, 3:, 3:] += make_psd(w * shape_hessian(q[b, 3:]))
    return G, H


def body_triplets(H):
    """10 upper 3x3 blocks per body: rows, cols, (m, 3, 3) values."""
    rows, cols, vals = [], [], []
    for b in range(H.shape[0]):
        for i in range(4):
            for j in range(i, 4):
                rows.append(4 * b + i); cols.append(4 * b + j)
                vals.append(H[b, 3 * i:3 * i + 3, 3 * j:3 * j + 3])
    return np.array(rows), np.array(cols), np.array(vals).reshape(-1, 3, 3)


def energy(q, qt, M, vol, fixed, kappa, dt):
    q = np.asarray(q, np.float64).reshape(-1, 12)
    qt = np.asarray(qt, np.float64).reshape(-1, 12)
    M = np.asarray(M, np.float64).reshape(-1, 12, 12)
    e = 0.0
    for b in range(q.shape[0]):
        if fixed is None or not fixed[b]:
            d = q[b] - qt[b]
            e += 0.5 * d @ M[b] @ d
        e += kappa * vol[b] * dt * dt * shape_energy(q[b, 3:])
    return e


def contact_gradient(g, body_of, xbar, nb, fixed=None):
    """(nb, 12) sum_i J_i^T g_i over the body vertices."""
    out = np.zeros((nb, 12))
    g = np.asarray(g, np.float64).reshape(-1, 3)
    for i in range(len(body_of)):
        b = body_of[i]
        if fixed is not None and fixed[b]:
            continue
        out[b] += jacobian(xbar[i]).T @ g[i]
    return out


def map_contacts(rows, cols, vals, nav, nb, body_of, xbar, fixed=None):
    """Vertex-space upper blocks -> DOF-space blocks (rows, cols, (m,3,3)).
    vals are (m, 3, 3) in row-major math orientation (vals[t][a, c])."""
    shift = 4 * nb - nav
    R, Cc, V = [], [], []
    fixed = np.zeros(nb, bool) if fixed is None else np.asarray(fixed, bool)
    for r, c, H in zip(rows, cols, vals):
        if not np.any(H):
            continue
        if r > c:
            r, c, H = c, r, H.T
        if r >= nav:
            R.append(r + shift); Cc.append(c + shift); V.append(H)
            continue
        bi = body_of[r]
        if fixed[bi] or (c < nav and fixed[body_of[c]]):
            continue
        Ji = jacobian(xbar[r])
        if c >= nav:
            X = Ji.T @ H                      # 12 x 3
            for k in range(4):
                R.append(4 * bi + k); Cc.append(c + shift); V.append(X[3 * k:3 * k + 3])
            continue
        bj = body_of[c]
        X = Ji.T @ H @ jacobian(xbar[c])      # 12 x 12
        if bi != bj:
            for k in range(4):
                for l in range(4):
                    R.append(4 * bi + k); Cc.append(4 * bj + l); V.append(X[3 * k:3 * k + 3, 3 * l:3 * l + 3])
            continue
        if r != c:
            X = X + X.T
        for k in range(4):
            for l in range(k, 4):
                R.append(4 * bi + k); Cc.append(4 * bi + l); V.append(X[3 * k:3 * k + 3, 3 * l:3 * l + 3])
    return np.array(R, np.int64), np.array(Cc, np.int64), np.array(V).reshape(-1, 3, 3)


def map_contacts_full_storage(rows, cols, vals, nav, nb, body_of, xbar, fixed=None):
    """The reference's procedure on full vertex-space storage (each entry
    (r, c, H) with r, c in any order): FEM-FEM entries shifted, lower ones
    dropped (zeroed in the reference); body-FEM entries (body row) -> the 4
    blocks of J_r^T H, FEM-body entries dropped (the lower class); body-body
    -> the 16 blocks of J_r^T H J_c, those below the diagonal dropped
    (written as zero blocks in the reference); fixed bodies -> nothing."""
    shift = 4 * nb - nav
    fixed = np.zeros(nb, bool) if fixed is None else np.asarray(fixed, bool)
    R, Cc, V = [], [], []
    for r, c, H in zip(rows, cols, vals):
        if r >= nav and c >= nav:
            if r <= c:
                R.append(r + shift); Cc.append(c + shift); V.append(H)
            continue
        if r >= nav:                          # FEM row, body column: lower class
            continue
        if fixed[body_of[r]] or (c < nav and fixed[body_of[c]]):
            continue
        Jr = jacobian(xbar[r])
        if c >= nav:
            X = Jr.T @ H
            for k in range(4):
                R.append(4 * body_of[r] + k); Cc.append(c + shift); V.append(X[3 * k:3 * k + 3])
            continue
        X = Jr.T @ H @ jacobian(xbar[c])
        for k in range(4):
            for l in range(4):
                row, col = 4 * body_of[r] + k, 4 * body_of[c] + l
                if row <= col:
                    R.append(row); Cc.append(col); V.append(X[3 * k:3 * k + 3, 3 * l:3 * l + 3])
    return np.array(R, np.int64), np.array(Cc, np.int64), np.array(V).reshape(-1, 3, 3)


def dense_upper(rows, cols, vals, nbr):
    """Dense symmetric matrix of upper-stored blocks (A = U + U^T - diag)."""
    A = np.zeros((3 * nbr, 3 * nbr))
    for r, c, B in zip(rows, cols, vals):
        A[3 * r:3 * r + 3, 3 * c:3 * c + 3] += B
        if r != c:
            A[3 * c:3 * c + 3, 3 * r:3 * r + 3] += B.T
    return A


def global_jacobian(nav, nf, nb, body_of, xbar):
    """Dense (3 (nav + nf)) x (12 nb + 3 nf) map from DOFs to vertex positions."""
    Jg = np.zeros((3 * (nav + nf), 12 * nb + 3 * nf))
    for i in range(nav):
        b = body_of[i]
        Jg[3 * i:3 * i + 3, 12 * b:12 * b + 12] = jacobian(xbar[i])
    Jg[3 * nav:, 12 * nb:] = np.eye(3 * nf)
    return Jg


def block_jacobi_inverse(A, nb):
    """12x12 inverses of the body diagonal blocks of a dense system."""
    return np.array([np.linalg.inv(A[12 * b:12 * b + 12, 12 * b:12 * b + 12]) for b in range(nb)])


def dyadic_mass(xbar, mass, vstart):
    """Per body sum_i m_i J_i^T J_i (the lumped form of ABDJacobiDyadicMass,
    abd_system/abd_jacobi_matrix.h:162-200)."""
    nb = len(vstart) - 1
    out = np.zeros((nb, 12, 12))
    for b in range(nb):
        for i in range(vstart[b], vstart[b + 1]):
            J = jacobian(xbar[i])
            out[b] += mass[i] * J.T @ J
    return out
