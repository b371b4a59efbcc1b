"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (gcc/g++
only, no GPU code): the FP64/FP32 oracle (every entry point, on a small grid
with exact-size heap buffers) and the executables' parameter, geometry and
binary I/O layer (fwi2d_amd/host/admin.cpp), including its error paths.
Any sanitizer report or leak fails the test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
HOST = os.path.join(ROOT, "fwi2d_amd", "host")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=undefined"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="1")

ORACLE_DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "fwi_oracle.h"
#define NEW(T, n) ((T *)calloc((size_t)(n), sizeof(T)))
int main(void)
{
    const int nz = 20, nx = 24, nb = 6, nzz = nz + 2 * nb, nxx = nx + 2 * nb, nt = 60;
    const long n = (long)nzz * nxx, m = (long)nz * nx;
    const double dh = 10.0, dt = 1e-3;
    double *w = NEW(double, nt), *re = NEW(double, nt / 2 + 1), *im = NEW(double, nt / 2 + 1);
    double *x2 = NEW(double, nt);
    const int tlag = or_wavelet(nt, dt, 30.0, w);
    or_dft_r2c(nt, w, re, im);
    or_dft_c2r(nt, re, im, x2);
    float *wf = NEW(float, nt);
    for (int k = 0; k < nt; ++k) wf[k] = (float)w[k];
    float *v = NEW(float, m), *vpad = NEW(float, n), *back = NEW(float, m), *vrb = NEW(float, n);
    for (long k = 0; k < m; ++k) v[k] = 1500.0f + (float)(k % nz) * 50.0f;
    or_expand(v, nz, nx, nb, vpad);
    or_reduce(vpad, nz, nx, nb, back);
    or_rb_velocity(vpad, nzz, nxx, nb, 0.0f, 0.4f, 1000.0f, 6000.0f, 7, 3, vrb);
    double *b1d = NEW(double, nb);
    or_abc_b1d(nb, 0.01, b1d);
    double d = or_damper(b1d, 0, 0, nzz, nxx, nb) + or_damper(b1d, nzz - 1, nxx - 1, nzz, nxx, nb);
    const int nrec = 5;
    int *riz = NEW(int, nrec), *rix = NEW(int, nrec);
    for (int r = 0; r < nrec; ++r) { riz[r] = nb + 1; rix[r] = nb + 2 + 4 * r; }
    const int siz = nb + 2, six = nb + nx / 2;
    double *seis = NEW(double, (long)nrec * nt), *P = NEW(double, n), *Q = NEW(double, n);
    float *seisf = NEW(float, (long)nrec * nt), *Pf = NEW(float, n), *Qf = NEW(float, n);
    or_forward_f64(nzz, nxx, nb, dh, dt, nt, tlag, vpad, w, siz, six, nrec, riz, rix, 1, b1d,
                   seis, P, Q, 1);
    or_forward_f32(nzz, nxx, nb, dh, dt, nt, tlag, vpad, wf, siz, six, nrec, riz, rix, 1, b1d,
                   seisf, Pf, Qf, 1);
    double *acc = NEW(double, n), *ill = NEW(double, n), *lap = NEW(double, n);
    float *accf = NEW(float, n), *illf = NEW(float, n), *lapf = NEW(float, n);
    for (int mode = 0; mode < 2; ++mode) {
        or_forward_f64(nzz, nxx, nb, dh, dt, nt, tlag, vrb, w, siz, six, nrec, riz, rix, 0, b1d,
                       NULL, P, Q, 1);
        or_reverse_f64(nzz, nxx, dh, dt, nt, tlag, vrb, P, Q, nrec, riz, rix, seis, mode, acc,
                       ill, 1);
        or_forward_f32(nzz, nxx, nb, dh, dt, nt, tlag, vrb, wf, siz, six, nrec, riz, rix, 0, b1d,
                       NULL, Pf, Qf, 1);
        or_reverse_f32(nzz, nxx, dh, dt, nt, tlag, vrb, Pf, Qf, nrec, riz, rix, seisf, mode,
                       accf, illf, 1);
    }
    or_laplacian_f64(nzz, nxx, dh, acc, lap);
    or_laplacian_f32(nzz, nxx, dh, accf, lapf);
    double *img = NEW(double, m), *il = NEW(double, m), *out = NEW(double, m);
    for (long k = 0; k < m; ++k) { img[k] = (double)(k % 7) - 3.0; il[k] = 1.0 + (double)(k % 5); }
    or_image_post(img, il, nz, nx, dh, out);
    or_gradient_post(img, il, nz, nx, dh, 40.0, out);
    const double a[3] = {0.0, 1.0, 2.0}, f[3] = {3.0, 1.0, 2.0};
    const double step = or_parabola_step(a, f);
    const int idx = or_grid_index(105.0f, (float)dh, nb);
    printf("ok %d %g %g %d %g\n", tlag, d, step, idx, seis[nt - 1]);
    free(w); free(re); free(im); free(x2); free(wf); free(v); free(vpad); free(back); free(vrb);
    free(b1d); free(riz); free(rix); free(seis); free(P); free(Q); free(seisf); free(Pf);
    free(Qf); free(acc); free(ill); free(lap); free(accf); free(illf); free(lapf); free(img);
    free(il); free(out);
    return 0;
}
"""


def build(cmd, tmp_path):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    return r


def clean_run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=300, **kw)
    for marker in ("AddressSanitizer", "LeakSanitizer", "runtime error"):
        assert marker not in r.stderr, r.stderr[-4000:]
    return r


def test_oracle_under_sanitizers(tmp_path):
    drv = tmp_path / "drv.c"
    drv.write_text(ORACLE_DRIVER)
    exe = tmp_path / "drv"
    srcs = [os.path.join(ORACLE, f) for f in
            ("oracle_f64.c", "oracle_f32.c", "oracle_misc.c", "oracle_rb.c")]
    build(["gcc", "-std=c99", "-fopenmp", *SAN, f"-I{ORACLE}", str(drv), *srcs, "-lm",
           "-o", str(exe)], tmp_path)
    r = clean_run([str(exe)])
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr[-2000:]


def test_host_io_under_sanitizers(tmp_path):
    from test_host_cpu import DRIVER, PARAMS
    drv = tmp_path / "drv.cpp"
    drv.write_text(DRIVER)
    exe = tmp_path / "drv"
    build(["g++", "-std=c++17", *SAN, f"-I{HOST}", str(drv), os.path.join(HOST, "admin.cpp"),
           "-o", str(exe)], tmp_path)
    p = tmp_path / "parameters.txt"
    p.write_text(PARAMS + "time_spacing = 8e-4\n")
    for key in ("x_samples", "model_file", "quoted", "listed", "tight", "empty", "nope"):
        clean_run([str(exe), "param", key, str(p)])
    clean_run([str(exe), "us", str(p)])
    sps, rps = tmp_path / "sps.txt", tmp_path / "rps.txt"
    sps.write_text("# x,z\n1000.00,100.00\n\n4000.00,100.00\n")
    rps.write_text("0.00,0.00\n2500.00,0.00\n5000.00,12.50\n")
    g = tmp_path / "g.txt"
    g.write_text(f"SPS = {sps}\nRPS = {rps}\n")
    assert clean_run([str(exe), "geom", str(g)]).returncode == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("1000.00\n")                      # malformed geometry line
    g.write_text(f"SPS = {bad}\nRPS = {rps}\n")
    assert clean_run([str(exe), "geom", str(g)]).returncode == 1
    b = tmp_path / "m.bin"
    b.write_bytes(bytes(40))
    assert clean_run([str(exe), "bin", str(b), "10"]).returncode == 0
    assert clean_run([str(exe), "bin", str(b), "11"]).returncode == 1
    assert clean_run([str(exe), "bin", str(tmp_path / "none.bin"), "1"]).returncode == 1
