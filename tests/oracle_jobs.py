"""Oracle twins of the jobs the GPU tests run: for a serving JobSpec or a
batched ChannelJob, the same initial state and the same boundary composite
built from the oracle (oracle/lbm_oracle.py), so that a device result can be
checked against oracle/lbm_oracle_native.run on the CPU.

Test helper only (not collected: the file name has no test_ prefix)."""

from __future__ import annotations

import numpy as np

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N


def spec_initial(spec):
    """(f0, rho0, u0) of JobSpec.start: equilibrium of rho = 1 and u = (u_in, 0)
    for a channel, of rest for a box."""
    if spec.kind == "von_karman":
        rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((spec.lx, spec.ly), spec.u_in)
    else:
        rho, u = O.density_1_velocity_0_initial((spec.lx, spec.ly))
    return O.equilibrium(rho, u), rho, u


def spec_boundary(spec):
    """The oracle composite of JobSpec.boundary()."""
    if spec.kind == "von_karman":
        return O.von_karman_boundary_conditions(spec.lx, spec.ly, 1.0, spec.u_in, spec.plate, walls=True,
                                                y_offset=1)
    if spec.kind == "couette":
        return O.couette_flow_boundary_conditions(spec.lx, spec.ly, spec.u_in, 1.0)
    if spec.kind == "poiseuille":
        drho = 3 * spec.delta_p
        return O.poiseuille_flow_boundary_conditions(spec.lx, spec.ly, (1 + drho / 2) / 3, (1 - drho / 2) / 3)
    raise ValueError(spec.kind)


def spec_probe(spec):
    return (3 * spec.lx // 4, spec.ly // 2) if spec.kind == "von_karman" else None


def run_spec(spec, steps=None, state=None):
    """Oracle run of a JobSpec (steps defaults to its budget); returns
    (f, rho, u, trace) with trace the (rho, ux, uy) rows at the probe."""
    f0, rho0, u0 = state if state is not None else spec_initial(spec)
    return N.run(f0, rho0, u0, spec.relaxation, spec.steps if steps is None else steps, spec_boundary(spec),
                 probe=spec_probe(spec))


def spec_result(spec, oracle_out):
    """The fields of JobSpec.result that the oracle's run determines: the
    probe samples of a channel, the u_x profile at x = lx/2 of a box."""
    f, rho, u, trace = oracle_out
    if spec.kind == "von_karman":
        return {"samples": trace}
    return {"ux_mid": u[spec.lx // 2, :, 0].copy()}


def channel_boundary(job):
    """Oracle composite of workloads.build_lattice(job)."""
    return O.von_karman_boundary_conditions(job.lx, job.ly, 1.0, job.u_in, job.plate, walls=True, y_offset=1)


def perturbed_channel_state(lx, ly, u_in, seed, scale=1e-4):
    """Equilibrium of rho = 1, u = (u_in, 0) with a seeded perturbation of every
    population, so that every node's arithmetic differs from its neighbours'."""
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), u_in)
    f = O.equilibrium(rho, u)
    f = f + scale * np.random.default_rng(seed).standard_normal(f.shape)
    return f, rho, u


def perturbed_rest_state(lx, ly, seed, scale=1e-4):
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    f = O.equilibrium(rho, u)
    f = f + scale * np.random.default_rng(seed).standard_normal(f.shape)
    return f, rho, u
