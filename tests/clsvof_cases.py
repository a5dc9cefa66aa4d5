"""Inputs of Clsvof_Scheme for the tests: a liquid bounded by a plane or a
sphere on a tensor grid, with exact cut-cell volume fractions, the signed
distance as the level set (negative in the liquid, as sign(0.5 - vof)
makes it) and a chosen velocity field."""

import numpy as np

from oracle import orc


def grid(shape, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), stretch=None, seed=0):
    rng = np.random.default_rng(seed)
    d = []
    for a, n in enumerate(shape):
        w = np.ones(n) if stretch is None or stretch[a] == 0 else 1.0 + stretch[a] * rng.random(n)
        d.append(w / w.sum() * (hi[a] - lo[a]))
    e = [lo[a] + np.concatenate([[0.0], np.cumsum(q)]) for a, q in enumerate(d)]
    c = [0.5 * (q[1:] + q[:-1]) for q in e]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    DX, DY, DZ = np.meshgrid(*d, indexing="ij")
    f = np.asfortranarray
    return dict(x=f(X), y=f(Y), z=f(Z), dx=f(DX), dy=f(DY), dz=f(DZ), edges=e)


def _halo(a, fill=0.0):
    I, J, K = a.shape
    h = np.full((I + 2, J + 2, K + 2), fill, order="F")
    h[1:-1, 1:-1, 1:-1] = a
    return h


def plane_liquid(g, n, t):
    """Liquid {n.x < t}: vof by cut volume, phi = n.x - t (the signed
    distance, negative in the liquid), normals n."""
    n = np.asarray(n, float) / np.linalg.norm(n)
    phi = n[0] * g["x"] + n[1] * g["y"] + n[2] * g["z"] - t
    vof = np.empty_like(phi)
    it = np.nditer(phi, flags=["multi_index"])
    for _ in it:
        i, j, k = it.multi_index
        d = (g["dx"][i, j, k], g["dy"][i, j, k], g["dz"][i, j, k])
        # liquid on the corner side of the plane through the cell: Find_Distance
        # convention s = corner distance of {n.x >= t}; liquid = V - that
        s = -phi[i, j, k] + 0.5 * (abs(n[0]) * d[0] + abs(n[1]) * d[1] + abs(n[2]) * d[2])
        vof[i, j, k] = orc.volume_fraction(*d, *n, s) / (d[0] * d[1] * d[2])
    return vof, phi, n


def state(g, vof, phi, n):
    I, J, K = vof.shape
    hv = _halo(vof)
    hp = _halo(phi)
    # ghosts as BoundaryConditionLvsVof leaves them: vof copied
    hv[0], hv[-1] = hv[1], hv[-2]
    hv[:, 0], hv[:, -1] = hv[:, 1], hv[:, -2]
    hv[:, :, 0], hv[:, :, -1] = hv[:, :, 1], hv[:, :, -2]
    nl = [np.asfortranarray(np.full((I + 2, J + 2, K + 2), n[a])) for a in range(3)]
    shapes = orc.clsvof_face_shapes((I, J, K))
    lvs = [np.asfortranarray(np.zeros(s)) for s in shapes]
    # the values the previous BC call leaves: phi at the face centre
    lvs[0] = np.asfortranarray(phi[0] - n[0] * g["dx"][0] / 2)
    lvs[2] = np.asfortranarray(phi[:, 0] - 0.0)
    lvs[4] = np.asfortranarray(phi[:, :, 0] - 0.0)
    return dict(vof_l=hv, phi_l=hp, nx_l=nl[0], ny_l=nl[1], nz_l=nl[2], lvs_var=lvs,
                vof_var=[np.asfortranarray(np.zeros(s)) for s in shapes])


def inputs(g, vel, dt, solid_n=(0.0, 0.0, 1.0), lvs_flag=(1,) * 6, lvs_const=(0.0,) * 6,
           vof_flag=(1,) * 6, vof_const=(0.0,) * 6):
    """vel = (u, v, w) halo arrays; BCu%VarW, BCv%VarS, BCw%VarB are the
    boundary face velocities u(0,:,:), v(:,0,:), w(:,:,0)."""
    u, v, w = (np.asfortranarray(a) for a in vel)
    I, J, K = g["dx"].shape
    hs = (I + 2, J + 2, K + 2)
    sn = [np.asfortranarray(np.full(hs, solid_n[a])) for a in range(3)]
    return dict(u=u, v=v, w=w, x=g["x"], y=g["y"], z=g["z"], dx=g["dx"], dy=g["dy"],
                dz=g["dz"], nx=sn[0], ny=sn[1], nz=sn[2],
                bcu_w=np.asfortranarray(u[0, 1:-1, 1:-1]), bcv_s=np.asfortranarray(v[1:-1, 0, 1:-1]),
                bcw_b=np.asfortranarray(w[1:-1, 1:-1, 0]), lvs_flag=list(lvs_flag),
                lvs_const=list(lvs_const), vof_flag=list(vof_flag), vof_const=list(vof_const),
                dt=dt)


def uniform_velocity(shape, U):
    I, J, K = shape
    hs = (I + 2, J + 2, K + 2)
    return [np.asfortranarray(np.full(hs, float(U[a]))) for a in range(3)]
