/*
 * Reference-compatible extern "C" entry points for the 2B/3B path (the Fortran / Python ctypes
 * ABI of MBX, src/bblock/external_call.cpp:47-427), implemented over bblock::System
 * (include/mbx/system.h) and the MI355X engine.  Same names, argument meaning and one global
 * system per process; existing callers switch by relinking against libmbxg.so.
 *
 * Energies cover the 1B + 2B + 3B terms; in get_energy_decomp_* the 4B, dispersion, Buckingham
 * and electrostatic outputs are written as 0 (those terms are outside this engine).  As in the
 * reference, an error (a C++ exception inside) ends the process: it is printed on stderr and
 * the process exits with EXIT_FAILURE, since these entry points have no error return.
 */
#ifndef MBX_EXTERNAL_CALL_H
#define MBX_EXTERNAL_CALL_H

#ifdef __cplusplus
extern "C" {
#endif

/* external_call.cpp:57 */
void initialize_system_(double* coords, int* nat_monomers, char at_names[][5], char monomers[][5], int* nmon,
                        char json_file[20]);
/* external_call.cpp:76: nranks independent systems of the same monomers (one per rank of the
 * caller), addressed by rank in the *2_ entry points */
void initialize_system2_(double* coords, int* nat_monomers, char at_names[][5], char monomers[][5], int* nmon,
                         char json_file[20], int* nranks);
/* external_call.cpp:106 */
void initialize_system_py_(double* coords, int* nat_monomers, char** at_names, char** monomers, int* nmon,
                           char* json_file);
/* external_call.cpp:131 */
void get_energy_(double* coords, int* nat, double* energy);
/* external_call.cpp:139 */
void get_energy2_(double* coords, int* nat, double* energy, int* rank);
/* external_call.cpp:208 */
void get_energy_g_(double* coords, int* nat, double* energy, double* grads);
/* external_call.cpp:219 */
void get_energy_g2_(double* coords, int* nat, double* energy, double* grads, int* rank);
/* external_call.cpp:237 */
void get_energy_pbc_(double* coords, int* nat, double* box, double* energy);
/* external_call.cpp:318 */
void get_energy_pbc_g_(double* coords, int* nat, double* box, double* energy, double* grads);
/* external_call.cpp:159 */
void get_energy_decomp_(double* coords, int* nat, double* E1b, double* E2b, double* E3b, double* E4b, double* Edisp,
                        double* Ebuck, double* Eelec);
/* external_call.cpp:261 */
void get_energy_decomp_pbc_(double* coords, int* nat, double* box, double* E1b, double* E2b, double* E3b,
                            double* E4b, double* Edisp, double* Ebuck, double* Eelec);
/* external_call.cpp:343 */
void get_virial_(double* virial);
/* external_call.cpp:355 */
void set_real_xyz_(double* coords, int* nat);
/* external_call.cpp:360: every site's coordinates (3 * GetNumSites doubles, input order; the
 * water M-site included) */
void get_xyz_(double* coords);
/* external_call.cpp:376 */
void set_box_(int* length, double* box);
/* external_call.cpp:422 */
void finalize_system_(void);
/* external_call.cpp:423 */
void finalize_system2_(void);

#ifdef __cplusplus
}
#endif
#endif
