/* vpfft_io.h -- input deck, microstructure files and initial microstructures
 * around the strain-step engine of vpfft.h.  Host-only: nothing here touches
 * the GPU, so these entry points also work on a machine without one.
 *
 * Errors: every call returns VPFFT_OK (0) or a negative VPFFT_E_* code from
 * vpfft.h, with the message in vpfft_last_error(). */
#ifndef VPFFT_IO_H_
#define VPFFT_IO_H_

#include <stdint.h>

#include "vpfft.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VPFFT_INIT_NONE = -1,
  VPFFT_INIT_TWO_GRAIN = 0, /* ConfigFile::eInitTwoGrain, Src/Config.h:24-28 */
  VPFFT_INIT_FILE = 1       /* ConfigFile::eInitFile */
};

/* Utilities::ConfigFile (Src/Config.h:21-52), the reference's input deck.
 * Keywords, one per line, '#' starts a comment line:
 *   GridDimension NumX NumY NumZ
 *   InputStrainRate            followed by three lines "s_i0 s_i1 s_i2"
 *   MaxNRIter n | TimeStep dt | NumTimeIter n | NRTolerence eps
 *   TwoGrainInit | FileInit <microstructure file>
 *   CRSS c_1 .. c_12 | GammaDotBase g_1 .. g_12 | RateSensitivity n_1 .. n_12
 * NRTolerence and TimeStep are stored rounded to float, as the reference's
 * members are `float` (Src/Config.h:39-40).  The per-slip lists need 12
 * values and fill slots 0..11 from tokens 1..12 (the reference asserts 11
 * and writes slots 1..12 of a 12-entry vector, Src/Config.cpp:190-215). */
typedef struct {
  int num_x, num_y, num_z;
  double input_strain_rate[3][3];
  int max_nr_iter;
  double nr_tolerance;
  double time_step;
  int num_time_iter;
  int init_method; /* VPFFT_INIT_* */
  char microstructure_file[1024];
  double crss[12];
  double gamma_dot_base[12];
  int rate_sensitivity[12];
} vpfft_deck;

/* ConfigFile::Read / ConfigFile::Parse (Src/Config.cpp:11-261).  Missing
 * required keywords, an unknown keyword, a short line or a missing FileInit
 * file return VPFFT_E_INVALID (the reference exits or returns false).  The
 * FileInit path is resolved relative to the deck's directory when it is not
 * absolute and does not exist as given. */
int vpfft_deck_read(const char* path, vpfft_deck* out);
int vpfft_deck_parse(const char* text, vpfft_deck* out);

/* The material a deck describes: its GammaDotBase and RateSensitivity, with
 * the reference's Voce defaults tau0 = tau1 = 11, theta0 = 430, theta1 = 110
 * and an all-ones hardening matrix (InitializeMemory, Src/MaterialGrid.cpp:322-336).
 * The deck's CRSS is the per-slip initial CRSS of every voxel. */
int vpfft_deck_material(const vpfft_deck* deck, vpfft_material* out);
/* EigenRep(InputStrainRate) (Src/LinearAlgebra.h:307-312): the deck's 3x3
 * strain rate as the 5-component deviatoric vector. */
int vpfft_deck_strain_rate(const vpfft_deck* deck, double edot_out[5]);

/* UnitTester::SetMaterialGrid(SMatrix3x3*, EigenRep*, ...)
 * (Src/UnitTester.cpp:46-78): two orientations in a quadrant checkerboard
 * over x and y -- Euler (-40, 0, 0) degrees where (i >= Nx/2) == (j >= Ny/2),
 * else (0, 40, 0) -- with the reference's stress placeholders +-1.
 * orientation_out: N x 9; stress_out: N x 5 (may be NULL). */
int vpfft_two_grain_init(int nx, int ny, int nz, double* orientation_out, double* stress_out);

/* SMatrix3x3::BuildActiveEulerMatrix (Src/LinearAlgebra.cpp:161-181) for
 * angles in degrees, with the reference's float-precision cosf / sinf and its
 * DEGREE_TO_RADIAN (Src/LinearAlgebra.h:21,28).  euler_deg: n x 3;
 * orientation_out: n x 9 row-major. */
int vpfft_euler_deg_to_matrix(const double* euler_deg, long n, double* orientation_out);

/* UnitTester::ReadRicardoData (Src/UnitTester.cpp:237-265): N lines of
 * "phi1 Phi phi2 x y z grain phase", angles in degrees, x y z 1-based, in
 * any order.  orientation_out: N x 9 (BuildActiveEulerMatrix as above);
 * euler_out (N x 3 degrees), grain_out, phase_out: may be NULL.  A voxel
 * outside the grid, a repeated or missing voxel, or a short file returns
 * VPFFT_E_INVALID. */
int vpfft_read_euler_file(const char* path, int nx, int ny, int nz, double* orientation_out, double* euler_out,
                          int32_t* grain_out, int32_t* phase_out);
/* The same format written (UnitTester::PrintGrid, Src/UnitTester.cpp:17-39):
 * z outermost, x innermost, 1-based indices.  grain / phase may be NULL
 * (written as 1). */
int vpfft_write_euler_file(const char* path, int nx, int ny, int nz, const double* euler_deg, const int32_t* grain,
                           const int32_t* phase);

/* UnitTester::WriteRandomMaterial / ReadRandomMaterial
 * (Src/UnitTester.cpp:116-168): per voxel in ToIndex order one line of the
 * 9 orientation entries (row-major) and the 5 stress components.  Written
 * with 17 significant digits, so a round trip is exact (the reference's
 * ostream default keeps 6).  stress may be NULL on write (zeros) and on read
 * (skipped). */
int vpfft_write_material_file(const char* path, int nx, int ny, int nz, const double* orientation,
                              const double* stress);
int vpfft_read_material_file(const char* path, int nx, int ny, int nz, double* orientation_out, double* stress_out);

/* A whole run driven by a deck, the intent of UnitTester::OpenMP_VPFFT_Test
 * (Src/UnitTester.cpp:327-367), which reads TestInput.config but then
 * hard-codes its grid and loading: builds the microstructure (TwoGrainInit,
 * or the FileInit file: 8 columns = Euler file, 14 = orientation + stress),
 * sets every voxel's CRSS to the deck's, voxel length 1, and calls
 * vpfft_run_vpfft(EigenRep(InputStrainRate), MaxNRIter, NumTimeIter,
 * TimeStep, NRTolerence).  stress_strain_path: the Stress_Strain.txt lines
 * (NULL: not written).  material_out_path: the final orientation and stress
 * in the material-file format (NULL: not written).  infos: per-step records,
 * up to infos_capacity (may be NULL). */
int vpfft_run_deck(const char* deck_path, int precision, const char* stress_strain_path,
                   const char* material_out_path, vpfft_step_info* infos, int infos_capacity);

#ifdef __cplusplus
}
#endif

#endif /* VPFFT_IO_H_ */
