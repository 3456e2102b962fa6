/*
 * ORACLE (test infrastructure only). CPU restatement of the serial FP64 plasma-density
 * recurrence that the reference runs as a CUDA RawKernel, one thread per (x, y) line:
 *   src/3+1D-UPPE-ERK4(3)IP-envelope/media_physics.py:11-66  (kernel text)
 *   src/3+1D-UPPE-ERK4(3)IP-envelope/media_physics.py:176-189 (launch: float32 scalars)
 *
 * Per line, for t = 0 .. nt-1 (rho starts at 0):
 *   rho_out[t]  = rho                      (value BEFORE the update)
 *   depl        = max((rho_nt - rho) / rho_nt, 0)
 *   gen_out[t]  = (w_pi * rho_nt + w_av * rho) * depl
 *   S           = w_pi * rho_nt * depl
 *   G           = w_av * depl - rate_rec
 *   gdt         = clamp(G * dt, -10, 10)
 *   f           = |gdt| < 1e-5 ? dt * (1 + gdt / 2) : (exp(gdt) - 1) / G
 *   rho         = clamp(rho * exp(gdt) + S * f, 0, rho_nt)
 * Inputs arrive as float32 (rates and scalars), all arithmetic is double.
 * Built by oracle/Makefile into oracle/_build/libplasma_k1.so and loaded with ctypes by
 * oracle/k1.py. Only tests, smoke() and the CPU-baseline leg of bench.py use it.
 */
#include <math.h>
#include <stddef.h>

void k1_plasma_lines(const float *w_ion, const float *w_ava, double *rho_out,
                     double *gen_out, float rate_rec_f, float dt_f, float rho_nt_f,
                     long nt, long nlines)
{
    const double rate_rec = (double)rate_rec_f;
    const double dt = (double)dt_f;
    const double rho_nt = (double)rho_nt_f;
    for (long line = 0; line < nlines; ++line) {
        const float *wi = w_ion + line * nt;
        const float *wa = w_ava + line * nt;
        double *ro = rho_out + line * nt;
        double *go = gen_out + line * nt;
        double rho = 0.0;
        for (long t = 0; t < nt; ++t) {
            ro[t] = rho;
            const double a = (double)wi[t];
            const double b = (double)wa[t];
            double depl = (rho_nt - rho) / rho_nt;
            if (depl < 0.0)
                depl = 0.0;
            go[t] = (a * rho_nt + b * rho) * depl;
            const double src = a * rho_nt * depl;
            const double g = b * depl - rate_rec;
            double gdt = g * dt;
            if (gdt > 10.0)
                gdt = 10.0;
            else if (gdt < -10.0)
                gdt = -10.0;
            const double e = exp(gdt);
            const double f = (fabs(gdt) < 1e-5) ? dt * (1.0 + 0.5 * gdt) : (e - 1.0) / g;
            rho = rho * e + src * f;
            if (rho > rho_nt)
                rho = rho_nt;
            if (rho < 0.0)
                rho = 0.0;
        }
    }
}
