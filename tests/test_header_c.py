"""The C ABI header is plain C.  A C11 program in the style of the integration
notes compiles against include/lbm_hip.h with gcc and links against the built
library (CPU tests); on a GPU it runs: it loads a channel (inlet, outlet and
plate, the reference's lattice_boltzmann_step with the closure of
milestoneQuickFunctionCalls.py:304-311), advances it with two steps per
launch, round-trips a checkpoint, reads |u| and the state, and everything
equals the CPU oracle bit for bit (FP64)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lattice_boltzmann_parallel_solver_amd")

PROGRAM = r"""
#include <stdio.h>
#include <stdlib.h>
#include "lbm_hip.h"

/* A channel: inlet (rho 1, u 0.1), outlet, plate mask; `steps` reference
 * steps from f = equilibrium(rho, u); writes f, rho, u and |u|. */
int run_channel(int nx, int ny, int storage, long steps, double omega, const uint8_t *plate, const double *rho,
                const double *u, double *f_out, double *rho_out, double *u_out, double *speed) {
    lbm_ctx *ctx;
    if (lbm_create(&ctx, nx, ny, storage, 0) != LBM_OK) {
        fprintf(stderr, "%s\n", lbm_last_error());
        return 1;
    }
    lbm_bc_desc d[3] = {{.kind = LBM_BC_INLET, .rho_in = 1.0, .u_in = 0.1},
                        {.kind = LBM_BC_OUTLET},
                        {.kind = LBM_BC_RIGID_OBJECT, .mask = plate}};
    lbm_bc_prog *prog;
    int rc = lbm_bc_compile(&prog, nx, ny, d, 3);
    if (rc == LBM_OK) rc = lbm_set_bc(ctx, prog);
    lbm_bc_destroy(prog);
    if (rc == LBM_OK) rc = lbm_set_state_equilibrium(ctx, rho, u, LBM_HOST);
    if (rc == LBM_OK) rc = lbm_set_temporal(ctx, 2, 0);
    if (rc == LBM_OK) rc = lbm_run(ctx, omega, steps);
    if (rc == LBM_OK) rc = lbm_speed(ctx, speed, LBM_HOST);
    int64_t n = 0;
    if (rc == LBM_OK) rc = lbm_checkpoint_bytes(ctx, &n);
    void *ck = malloc((size_t)n);
    if (rc == LBM_OK) rc = lbm_checkpoint_save(ctx, ck, n);
    if (rc == LBM_OK) rc = lbm_checkpoint_load(ctx, ck, n);
    free(ck);
    if (rc == LBM_OK) rc = lbm_get_state(ctx, f_out, rho_out, u_out, LBM_HOST);
    if (rc != LBM_OK) fprintf(stderr, "%s\n", lbm_last_error());
    lbm_destroy(ctx);
    return rc == LBM_OK ? 0 : 1;
}

static int io(const char *path, void *buf, size_t bytes, int write) {
    FILE *fp = fopen(path, write ? "wb" : "rb");
    if (!fp) return 1;
    size_t k = write ? fwrite(buf, 1, bytes, fp) : fread(buf, 1, bytes, fp);
    fclose(fp);
    return k != bytes;
}

/* use_lbm NX NY STEPS OMEGA DIR: reads DIR/plate.u8, rho.f64, u.f64; writes
 * DIR/f.out, rho.out, u.out, speed.out.  Without arguments: ABI version check. */
int main(int argc, char **argv) {
    if (argc < 6) return lbm_version() > 0 ? 0 : 1;
    int nx = atoi(argv[1]), ny = atoi(argv[2]);
    long steps = atol(argv[3]);
    double omega = atof(argv[4]);
    size_t n = (size_t)nx * ny;
    char path[4096];
    uint8_t *plate = malloc(n);
    double *rho = malloc(8 * n), *u = malloc(16 * n), *f = malloc(72 * n), *ro = malloc(8 * n), *uo = malloc(16 * n),
           *sp = malloc(8 * n);
    int bad = 0;
    snprintf(path, sizeof path, "%s/plate.u8", argv[5]);
    bad |= io(path, plate, n, 0);
    snprintf(path, sizeof path, "%s/rho.f64", argv[5]);
    bad |= io(path, rho, 8 * n, 0);
    snprintf(path, sizeof path, "%s/u.f64", argv[5]);
    bad |= io(path, u, 16 * n, 0);
    if (bad) return 2;
    if (run_channel(nx, ny, LBM_STORE_F64, steps, omega, plate, rho, u, f, ro, uo, sp)) return 3;
    snprintf(path, sizeof path, "%s/f.out", argv[5]);
    bad |= io(path, f, 72 * n, 1);
    snprintf(path, sizeof path, "%s/rho.out", argv[5]);
    bad |= io(path, ro, 8 * n, 1);
    snprintf(path, sizeof path, "%s/u.out", argv[5]);
    bad |= io(path, uo, 16 * n, 1);
    snprintf(path, sizeof path, "%s/speed.out", argv[5]);
    bad |= io(path, sp, 8 * n, 1);
    return bad ? 4 : 0;
}
"""

LIB = os.path.join(PKG, "liblbm_hip.so")


@pytest.fixture(scope="module")
def c_source(tmp_path_factory):
    d = tmp_path_factory.mktemp("capi_c")
    src = d / "use_lbm.c"
    src.write_text(PROGRAM)
    return d, src


def _link(d, src):
    return subprocess.run(["gcc", "-std=c11", str(src), "-I", os.path.join(ROOT, "include"), "-L", PKG, "-llbm_hip",
                           f"-Wl,-rpath,{PKG}", "-o", str(d / "use_lbm")], capture_output=True, text=True)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_header_compiles_as_c11(c_source):
    d, src = c_source
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-c", str(src), "-I", os.path.join(ROOT, "include"),
                        "-o", str(d / "use_lbm.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_c_program_links_against_the_library(c_source):
    if not os.path.exists(LIB):
        pytest.skip("liblbm_hip.so not built")
    d, src = c_source
    r = _link(d, src)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_c_program_runs_a_channel_equal_to_the_oracle(c_source):
    from oracle import lbm_oracle as O
    from oracle import lbm_oracle_native as N

    d, src = c_source
    r = _link(d, src)
    assert r.returncode == 0, r.stderr
    nx, ny, plate_size, steps = 96, 40, 10, 37
    omega = 1.0 / (3 * 0.04 + 0.5)
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((nx, ny), 0.1)
    plate = O.plate_mask(nx, ny, plate_size)
    plate.astype(np.uint8).tofile(str(d / "plate.u8"))
    rho.tofile(str(d / "rho.f64"))
    u.tofile(str(d / "u.f64"))
    print(f"case: C program, {nx}x{ny} channel, {steps} FP64 steps, two per launch", flush=True)
    run = subprocess.run([str(d / "use_lbm"), str(nx), str(ny), str(steps), repr(omega), str(d)],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    f = np.fromfile(str(d / "f.out")).reshape(nx, ny, 9)
    ro = np.fromfile(str(d / "rho.out")).reshape(nx, ny)
    uo = np.fromfile(str(d / "u.out")).reshape(nx, ny, 2)
    sp = np.fromfile(str(d / "speed.out")).reshape(nx, ny)
    bc = O.von_karman_boundary_conditions(nx, ny, 1.0, 0.1, plate_size)
    fw, rw, uw, _ = N.run(O.equilibrium(rho, u), rho, u, omega, steps, bc)
    assert np.array_equal(f, fw)
    assert np.array_equal(ro, rw)
    assert np.array_equal(uo, uw)
    assert np.array_equal(sp, np.sqrt(uw[..., 0] * uw[..., 0] + uw[..., 1] * uw[..., 1]))
