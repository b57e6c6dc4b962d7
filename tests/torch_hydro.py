"""TEST INFRASTRUCTURE: the 2-D hydro step restated in torch float64 (CPU) so
that autograd gives an FP64 reference gradient of the discrete map for the
adjoint tests. Follows the same formulas as the oracle (Eos.hpp:54-131,
Reconstruct.hpp:217-274, Riemann.hpp:22-176, TimeStepping.cpp SSPRK3, periodic
Boundaries.hpp fill) vectorised over the grid. Branch conventions: min/max
and the HLLC switches use torch.where on the forward values (min/max keep the
first operand on ties), |x| differentiates to sign(x) with sign(0) = 0.
"""
import torch

TINY = 1e-20


def _smin(a, b):
    return torch.where(b < a, b, a)


def _smax(a, b):
    return torch.where(a < b, b, a)


def cons_to_prim(q, g):
    rho = q[0]
    vx = q[1] / rho
    vy = q[2] / rho
    p = (g - 1.0) * (q[3] - 0.5 * rho * (vx * vx + vy * vy))
    return torch.stack([rho, vx, vy, p])


def prim_to_cons(w, g):
    rho, vx, vy, p = w
    return torch.stack([rho, rho * vx, rho * vy, p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy)])


def prim_to_flux(w, g, axis):
    rho, vx, vy, p = w
    vn = vx if axis == 0 else vy
    mf = rho * vn
    e = p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy)
    f1 = mf * vx + (p if axis == 0 else 0.0)
    f2 = mf * vy + (p if axis == 1 else 0.0)
    return torch.stack([mf, f1, f2, (e + p) * vn])


def weno5z(s):
    """s: list of 5 tensors (s-2..s2) -> (wL, wR)."""
    sm2, sm1, s0, sp1, sp2 = s
    c0, c1, eps = 13.0 / 12.0, 0.25, 1e-36
    b0 = c0 * (sm2 - 2 * sm1 + s0) ** 2 + c1 * (sm2 - 4 * sm1 + 3 * s0) ** 2
    b1 = c0 * (sm1 - 2 * s0 + sp1) ** 2 + c1 * (sm1 - sp1) ** 2
    b2 = c0 * (s0 - 2 * sp1 + sp2) ** 2 + c1 * (3 * s0 - 4 * sp1 + sp2) ** 2
    tau = torch.abs(b0 - b2)
    q0 = 1 + (tau / (b0 + eps)) ** 2
    q1 = 1 + (tau / (b1 + eps)) ** 2
    q2 = 1 + (tau / (b2 + eps)) ** 2
    eL = (2 * s0 + 5 * sm1 - sm2, -sp1 + 5 * s0 + 2 * sm1, 2 * sp2 - 7 * sp1 + 11 * s0)
    eR = (2 * s0 + 5 * sp1 - sp2, -sm1 + 5 * s0 + 2 * sp1, 2 * sm2 - 7 * sm1 + 11 * s0)
    aL = (3 * q0, 6 * q1, q2)
    aR = (3 * q2, 6 * q1, q0)
    wl = (eL[0] * aL[0] + eL[1] * aL[1] + eL[2] * aL[2]) / (6 * (aL[0] + aL[1] + aL[2]))
    wr = (eR[0] * aR[0] + eR[1] * aR[1] + eR[2] * aR[2]) / (6 * (aR[0] + aR[1] + aR[2]))
    return wl, wr


def rusanov(wl, wr, g, axis):
    iv = 1 + axis
    ql, qr = prim_to_cons(wl, g), prim_to_cons(wr, g)
    fl, fr = prim_to_flux(wl, g, axis), prim_to_flux(wr, g, axis)
    csl = torch.sqrt(g * wl[3] / wl[0])
    csr = torch.sqrt(g * wr[3] / wr[0])
    c = 0.5 * (csl + torch.abs(wl[iv]) + csr + torch.abs(wr[iv]))
    return 0.5 * (fl + fr - c * (qr - ql))


def hllc(wl, wr, g, axis):
    iv = 1 + axis
    im = iv
    csl = torch.sqrt(g * wl[3] / wl[0])
    csr = torch.sqrt(g * wr[3] / wr[0])
    ql, qr = prim_to_cons(wl, g), prim_to_cons(wr, g)
    pl, pr = wl[3], wr[3]
    rho_avg = 0.5 * (wl[0] + wr[0])
    cs_avg = 0.5 * (csl + csr)
    p_mid = 0.5 * (pl + pr + (wl[iv] - wr[iv]) * rho_avg * cs_avg)
    gf = (g + 1.0) / (2.0 * g)
    cls = torch.where(p_mid <= pl, torch.ones_like(pl), torch.sqrt(torch.clamp(1 + gf * (p_mid / pl - 1), min=0.0)))
    crs = torch.where(p_mid <= pr, torch.ones_like(pr), torch.sqrt(torch.clamp(1 + gf * (p_mid / pr - 1), min=0.0)))
    al = wl[iv] - csl * cls
    ar = wr[iv] + csr * crs
    bm = _smin(torch.full_like(al, -TINY), al)
    bp = _smax(torch.full_like(ar, TINY), ar)
    ul, ur = wl[iv] - al, wr[iv] - ar
    tl = pl + ul * ql[im]
    tr = pr + ur * qr[im]
    ml = wl[0] * ul
    mr = -(wr[0] * ur)
    am = (tl - tr) / (ml + mr)
    cp = _smax((ml * tr + mr * tl) / (ml + mr), torch.zeros_like(am))
    ul2, ur2 = wl[iv] - bm, wr[iv] - bp
    mfl, mfr = wl[0] * ul2, wr[0] * ur2

    def side(w, q, mf, u):
        f = [mf, mf * w[1], mf * w[2], q[3] * u + w[3] * w[iv]]
        f[im] = f[im] + w[3]
        return f

    fl = side(wl, ql, mfl, ul2)
    fr = side(wr, qr, mfr, ur2)
    pos = am >= 0
    sl = torch.where(pos, am / (am - bm), torch.zeros_like(am))
    sr = torch.where(pos, torch.zeros_like(am), -am / (bp - am))
    sm = torch.where(pos, -bm / (am - bm), bp / (bp - am))
    f = [sl * fl[k] + sr * fr[k] for k in range(4)]
    f[im] = f[im] + sm * cp
    f[3] = f[3] + sm * cp * am
    return torch.stack(f)


def fill_periodic(q, ng):
    """q: [4, yc, xc] with interior [ng:-ng, ng:-ng] -> ghosts filled (x then y)."""
    inner = q[:, ng:-ng, ng:-ng]
    x = torch.cat([inner[:, :, -ng:], inner, inner[:, :, :ng]], dim=2)
    return torch.cat([x[:, -ng:, :], x, x[:, :ng, :]], dim=1)


def stage_L(q, g, dtdx, ng, solver):
    """(dt/dx) * flux divergence on the interior of the ghost-filled q ([4, yc, xc])."""
    w = cons_to_prim(q, g)
    ny, nx = q.shape[1] - 2 * ng, q.shape[2] - 2 * ng
    riem = rusanov if solver == "rusanov" else hllc

    def sweep(axis):
        # reconstruct cells ng-1 .. ng+n along the axis, interior across
        if axis == 0:
            cells = [w[:, ng:ng + ny, ng - 1 + o - 2: ng + nx + 1 + o - 2] for o in range(5)]
        else:
            cells = [w[:, ng - 1 + o - 2: ng + ny + 1 + o - 2, ng:ng + nx] for o in range(5)]
        wl, wr = weno5z(cells)
        if axis == 0:
            left, right = wr[:, :, :-1], wl[:, :, 1:]  # faces ng .. ng+nx
        else:
            left, right = wr[:, :-1, :], wl[:, 1:, :]
        f = riem(left, right, g, axis)
        return (f[:, :, :-1] - f[:, :, 1:]) if axis == 0 else (f[:, :-1, :] - f[:, 1:, :])

    return dtdx * (sweep(0) + sweep(1))


def gravity_source(q, grav, dt, ng):
    """S*dt of the gravity term (Gravity.cpp:22-38) on the interior of q."""
    inner = q[:, ng:-ng, ng:-ng]
    zero = torch.zeros_like(inner[0])
    return dt * torch.stack([zero, inner[0] * grav[0], inner[0] * grav[1],
                             inner[1] * grav[0] + inner[2] * grav[1]])


def ssprk3_step(q, g, dt, dx, ng, solver, grav=None):
    dtdx = dt / dx

    def with_interior(qbase, new_inner):
        out = qbase.clone()
        out[:, ng:-ng, ng:-ng] = new_inner
        return fill_periodic(out, ng)

    inner = lambda a: a[:, ng:-ng, ng:-ng]  # noqa: E731

    def rhs(a):
        r = stage_L(a, g, dtdx, ng, solver)
        return r if grav is None else r + gravity_source(a, grav, dt, ng)

    q1 = with_interior(q, inner(q) + rhs(q))
    q2 = with_interior(q, 0.75 * inner(q) + 0.25 * (inner(q1) + rhs(q1)))
    q3 = with_interior(q, (1.0 / 3.0) * inner(q) + (2.0 / 3.0) * (inner(q2) + rhs(q2)))
    return q3
