/* pearl_input.h — C ABI of the drop-in program's input-file reader
 * (libpearl_input.so, host code only: no device, no HDF5).
 *
 * The reference reads its input inside the program (read_config,
 * pearl_serial.cpp:1600-1758, with the temperature laws :186-276 and the
 * FILL admission of parse_filling_operation :1209-1245); it has no library
 * entry for it.  These entry points expose the one reader the drop-in
 * program uses (csrc/host/settings.hpp), so that other hosts (the Python
 * binding, test harnesses) read input files with the same code, defaults,
 * checks and error texts instead of a second parser.
 *
 * Conventions: plain pointers and sizes; on failure an entry point returns
 * -1 and copies the reference's exception text into err (NUL-terminated,
 * truncated to err_len).
 */
#ifndef PEARL_INPUT_H
#define PEARL_INPUT_H

#include "pearl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The run settings of an input file (read_config :1600-1758) as one JSON
 * object: every setting under its input key's name (the derived ones too:
 * thermo, the effective coefficients at the run temperature; initial_mu;
 * the defaults the reference computes), reals as C99 hexadecimal strings
 * ("%a", exact), the FILL lines in file order, and the physics signature
 * stored in checkpoints (:1566-1597).  Writes the text when json_len exceeds
 * its length and returns the length (without the NUL) in every successful
 * case, so a caller can size its buffer and call again; -1 on an invalid
 * input. */
long long pearl_read_input(const char* path, char* json, long long json_len, char* err, int err_len);

/* The sharp labels of an input as FILL operations, the form
 * pearl_fill_phases takes: its admitted FILL lines for initialization =
 * microsim (:1772-1776), the two boxes of sharp_smooth (:1777-1786).  Writes
 * min(count, max_ops) records and the count; -1 when the input is invalid
 * or has no sharp labels. */
int pearl_input_fills(const char* path, pearl_fill* ops, int max_ops, int* count, char* err, int err_len);

/* The temperature laws (:203-218 linear_gp, :236-276
 * phase_diagram_fixed_curvature) in the reference's long-double
 * expressions.  Parabolas are {A, B, C, D}; dg_dT = {dq2/dT, dq1/dT,
 * dq0/dT}; c_eq = alpha, theta, gamma contacts at the eutectoid; slope =
 * alpha on a/g, gamma on a/g, theta on t/g, gamma on t/g.  0 or -1. */
int pearl_linear_gp_law(const double at_eutectoid[4], const double dg_dT[3], double undercooling, double out[4],
                        char* err, int err_len);
int pearl_phase_diagram_law(const double at_eutectoid[12], const double c_eq[3], const double slope[4],
                            double undercooling, double out[12], char* err, int err_len);

#ifdef __cplusplus
}
#endif

#endif /* PEARL_INPUT_H */
