/*
 * chns_ref.h - the reference's time-loop entry points, same names and
 * argument lists, over the C ABI of chns.h (libchns_ref.so).
 *
 * The reference (phynadia/Ternary-CHNS) declares its hot-path functions as
 * free void functions with K&R prototypes (global.h:79-98,156-160) and calls
 * them from main.c:276-284:
 *
 *     etd2rk(cwz, cphi1, cphi2);  free_energy(cphi1, cphi2);
 *     kinetic_energy(cwz, cphi1, cphi2);  perimeter(cphi1, cphi2);
 *     if (it % (navg/2) == 0) spectrum(cwz, cphi1, cphi2);
 *
 * This header gives those names full prototypes with the TERNARY argument
 * lists; a BINARY build passes its cphi as cphi1 and NULL for cphi2, an NS
 * build NULL for both.  `double (*)[2]` is the layout of FFTW's
 * `fftw_complex *` (the IDXC arrays [Nx][Ny/2+1], global.h:13), so the
 * reference's arrays pass unchanged and no FFTW header is needed.
 *
 * One process-wide context holds the state on the GPU.  With mirror != 0 at
 * chns_ref_init the host arrays are the state, as in the reference: each of
 * the calls below uploads the arrays it is given, runs, and (etd2rk,
 * spectrum) writes the state back into them in place.  With mirror == 0 the
 * state stays on the device between calls, the array arguments are not
 * touched, and the host arrays are read and written only by chns_ref_push
 * and chns_ref_pull (e.g. before main.c's own dump code reads them).  The
 * two modes agree to round-off (each upload makes the next step rebuild its
 * first inverse transforms, from 256^2 up with another radix plan than the
 * fused rebuild); the second moves no state over PCIe.
 *
 * Errors: the reference returns none and aborts through MPI_Abort; these
 * functions print the library's message to stderr and exit(1), and
 * free_energy exits with status 3 when the free-energy guard of
 * free_energy.c:151-154 (ephi >= 100) or the device's non-finite flag trips,
 * after writing its line.
 */
#ifndef CHNS_REF_H
#define CHNS_REF_H

#include <stdio.h>

#include "chns.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces main.c:37-194 (MPI and FFTW set-up, allocation, plans, kxky()):
 * the process-wide context for parameters p on HIP device `device`.
 * Returns a CHNS_* code (the only call here that returns one). */
int chns_ref_init(const chns_params *p, int device, int mirror);
/* Replaces main.c:380-413 (frees and MPI_Finalize). */
void chns_ref_finalize(void);

/* The reference's open output files (main.c:100-165): fu data/ken_t.dat,
 * fe data/free_en_t.dat, fdef1/fdef2 data/gamma1.dat, gamma2.dat (binary:
 * fdef1 = data/gamma.dat).  NULL for a file the model has not. */
void chns_ref_outputs(FILE *fu, FILE *fe, FILE *fdef1, FILE *fdef2);
/* The time t the diagnostics print (the reference's global t, advanced at
 * main.c:263 before each step). */
void chns_ref_set_time(double t);

/* Host state -> device (after init(), or after the caller changed it). */
void chns_ref_push(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]);
/* Device state -> host arrays (NULL skips a field). */
void chns_ref_pull(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]);
/* init.c:253-262,291-294: packed real [Nx][Ny] fields (wz NULL = zero
 * vorticity) -> dfft + truncation on the device; the resulting state is also
 * written to the host arrays given (NULL skips). */
void chns_ref_init_real(const double *wz, const double *phi1, const double *phi2,
                        double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]);

/* etd2rk.c:14-224: one ETD2RK step of the state. */
void etd2rk(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]);
/* free_energy.c:13-157: appends "t ephi egphi" to fe (:149); guard :151-154. */
void free_energy(double (*cphi1)[2], double (*cphi2)[2]);
/* kinetic_energy.c:3-144: appends "t eu [eph1 [eph2]]" to fu (:109,135,141). */
void kinetic_energy(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]);
/* perimeter.c:127-266: appends "t Sp_0 Sp_t def" per phase (:123,264-265). */
void perimeter(double (*cphi1)[2], double (*cphi2)[2]);
/* spectrum.c:3-194: writes spectra/{spectrum,Xspectrum,Yspectrum}.N
 * (N counts the calls from 0) and truncates the phase fields in place
 * (:22-27). */
void spectrum(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]);
/* The DUMP_REAL_MPI block of main.c:301-360 (truncate != 0) or the t = 0
 * dump of main.c:230-250 (truncate == 0): vort/wz.N and phase_field/phi*.N as
 * raw float64 [Nx][Ny] (write_array.c:48-79), N counting the calls from 0. */
void chns_ref_dump(int truncate);

#ifdef __cplusplus
}
#endif
#endif /* CHNS_REF_H */
