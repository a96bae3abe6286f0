"""2D acoustic wave propagation with C-PML absorbing boundaries.

Restates seismicCPML2D (traditional_solutions/seismic_cpml/
seismic_CPML_2D_pressure_second_order.py:12-150) and its damping profiles
(seismic_CPML_helper.py:12-88): second-order staggered finite differences for
the pressure p with kappa = rho c^2,
    p^{n+1} = 2 p^n - p^{n-1} + dt^2 kappa (d/dx(1/rho dp/dx) + d/dy(1/rho dp/dy)),
with the C-PML memory variables on both derivative stages and Dirichlet
p = 0 on the outer edge.  Returns the wavefields of every step.
"""

import numpy as np


def damping_profiles(velocity, npml, rcoef, k_max, alpha_max, npower, dt, deltas, dtype=np.float64):
    """Per-axis (a, a_half, b, b_half, K, K_half) C-PML profiles, shaped to
    broadcast against `velocity`."""
    profiles = []
    maxvel = np.max(velocity)
    for axis, N in enumerate(velocity.shape):
        delta = deltas[axis]
        thick = npml * delta
        d0 = -(npower + 1) * maxvel * np.log(rcoef) / (2 * thick)
        val = delta * np.arange(N)
        out = {}
        for tag, pos in (("", val), ("_half", val + delta / 2.0)):
            d = np.zeros(N)
            K = np.ones(N)
            alpha = np.zeros(N)
            for absc in (thick - pos, pos - ((N - 1) * delta - thick)):
                inside = absc >= 0.0
                xn = np.where(inside, absc / thick, 0.0)
                d = np.where(inside, d0 * xn ** npower, d)
                K = np.where(inside, 1 + (k_max - 1) * xn ** npower, K)
                alpha = np.where(inside, alpha_max * (1 - xn), alpha)
            b = np.exp(-(d / K + alpha) * dt)
            a = np.zeros(N)
            nz = np.abs(d) > 1e-6
            a[nz] = d[nz] * (b[nz] - 1) / (K[nz] * (d[nz] + K[nz] * alpha[nz]))
            shape = [1] * velocity.ndim
            shape[axis] = N
            out["a" + tag] = a.reshape(shape).astype(dtype)
            out["b" + tag] = b.reshape(shape).astype(dtype)
            out["K" + tag] = K.reshape(shape).astype(dtype)
        profiles.append(out)
    return profiles


def seismicCPML2D(NX, NY, NSTEPS, DELTAX, DELTAY, DELTAT, NPOINTS_PML, velocity, density, initial_pressures,
                  f0=20.0, dtype=np.float32, output_wavefields=True, gather_is=None):
    velocity = np.asarray(velocity).astype(dtype)
    density = np.asarray(density).astype(dtype)
    courant = np.max(velocity) * DELTAT * np.sqrt(1 / DELTAX ** 2 + 1 / DELTAY ** 2)
    if courant > 1.0:
        raise Exception("ERROR: time step is too large, simulation will be unstable %.2f" % courant)
    px, py = damping_profiles(velocity, NPOINTS_PML, 0.001, 1.0, 2.0 * np.pi * (f0 / 2.0), 2.0, DELTAT,
                              (DELTAX, DELTAY), dtype)
    kappa = density * velocity ** 2
    p_now = np.asarray(initial_pressures[1]).astype(dtype)
    p_old = np.asarray(initial_pressures[0]).astype(dtype)
    mdx = np.zeros((NX, NY), dtype)
    mdy = np.zeros((NX, NY), dtype)
    mdxx = np.zeros((NX, NY), dtype)
    mdyy = np.zeros((NX, NY), dtype)
    wavefields = np.zeros((NSTEPS, NX, NY), dtype) if output_wavefields else None
    gather = np.zeros((gather_is.shape[0], NSTEPS), dtype) if gather_is is not None else None
    rho_hx = np.pad(0.5 * (density[1:, :] + density[:-1, :]), [[0, 1], [0, 0]], mode="edge")
    rho_hy = np.pad(0.5 * (density[:, 1:] + density[:, :-1]), [[0, 0], [0, 1]], mode="edge")
    for it in range(NSTEPS):
        vdx = np.pad((p_now[1:, :] - p_now[:-1, :]) / DELTAX, [[0, 1], [0, 0]])
        vdy = np.pad((p_now[:, 1:] - p_now[:, :-1]) / DELTAY, [[0, 0], [0, 1]])
        mdx = px["b_half"] * mdx + px["a_half"] * vdx
        mdy = py["b_half"] * mdy + py["a_half"] * vdy
        pxx = (vdx / px["K_half"] + mdx) / rho_hx
        pyy = (vdy / py["K_half"] + mdy) / rho_hy
        vdxx = np.pad((pxx[1:, :] - pxx[:-1, :]) / DELTAX, [[1, 0], [0, 0]])
        vdyy = np.pad((pyy[:, 1:] - pyy[:, :-1]) / DELTAY, [[0, 0], [1, 0]])
        mdxx = px["b"] * mdxx + px["a"] * vdxx
        mdyy = py["b"] * mdyy + py["a"] * vdyy
        lap = (vdxx / px["K"] + mdxx) + (vdyy / py["K"] + mdyy)
        p_new = -p_old + 2 * p_now + DELTAT * DELTAT * lap * kappa
        p_new[0, :] = p_new[-1, :] = 0.0
        p_new[:, 0] = p_new[:, -1] = 0.0
        if wavefields is not None:
            wavefields[it] = p_now
        if gather is not None:
            gather[:, it] = p_now[gather_is[:, 0], gather_is[:, 1]]
        if np.max(np.abs(p_now)) > 1e25:
            raise Exception("code became unstable and blew up")
        p_old, p_now = p_now, p_new
    return [wavefields, gather]
