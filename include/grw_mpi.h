/*
 * grw_mpi.h -- the reference's MPI-layer entry points, exported with their
 * own names and signatures by libgrw_mpi.so and backed by one device context
 * of libgrw_fru.so (include/grw_fru.h).
 *
 * A program built from the reference's driver and global model (caf.c,
 * rk_pd54m.c, fcn.c, initialize_*.c, data_io.c, write_fru_props.c,
 * dynamicFRU.c, parameters_fcn_fru.c, mt19937_local.c) links against this
 * library instead of mpi_master.o and mpi_slave.o, and against no MPI at all:
 * the release units and their MT19937 streams live on one GPU, and every
 * call below is one call of include/grw_fru.h.  The FRU kernels
 * (simfru_local.c, fru_rates_local.c, fcn_fru.c, calc_fru_*_local.c) are not
 * linked either.
 *
 * The functions read the program's globals NFRU_sim and NFRU_scale (caf.c:
 * 34-36) as mpi_master.c does.  Arrays are the reference's: rows of
 * NFRU_sim_max FRUs, of which the first NFRU_sim are used, MT words in
 * unsigned long (LP64).  The reference's calls return nothing; on a failure of
 * the device library these print "grw_mpi: <call>: <reason>" to stderr and
 * exit the program with status 1.
 *
 * Environment (read at initialize_mpi_state):
 *   GRW_PRECISION=mixed   FP32 rates and local-Ca right-hand side (default fp64)
 *   GRW_SEQUENTIAL_SUMS=1 flux and average sums in the reference's sequential
 *                         FRU order (calc_fru_flux_local.c:45-70), for runs
 *                         compared bit for bit with the reference (default:
 *                         the device's fixed-order tree)
 *   GRW_DEVICE=<n>        HIP device ordinal (default 0)
 */
#ifndef GRW_MPI_H
#define GRW_MPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* mpi_master.c:625-635.  No processes: *rank = 0, *procs = 1. */
void initialize_mpi(int *argc, char **argv, int *rank, int *procs);
/* mpi_master.c:643-651.  Destroys the device context. */
void end_mpi(void);
/* mpi_slave.c:348 (caf.c:281 calls it on ranks other than 0; with one
 * process it is never reached).  Returns at once. */
void parallel(void);

/* mpi_master.c:199-361.  (Re)creates the device context for NFRU_sim FRUs at
 * NFRU_scale when either changed, and uploads the first NFRU_sim rows. */
void initialize_mpi_state(double FRU_states[][5], int LType_state[][4][2], int RyR_state[][4][5],
                          int Ito2_state[][4], int mti[], unsigned long mt[][625]);
/* mpi_master.c:509-617.  Downloads the first NFRU_sim rows. */
void parallel_get_FRUs(double FRU_states[][5], int LType_state[][4][2], int RyR_state[][4][5],
                       int Ito2_state[][4], unsigned long mt[][625], int mti[]);

/* mpi_master.c:369-469 (grw_distrib_simfru). */
void distrib_simFRU(double st_time, double end_time, double FRUdep_states0[3],
                    double FRUdep_statesf[3], double *Jxfer, double *Jtr, double *ICa,
                    double *Ito2);
/* mpi_master.c:41-113 (grw_calc_flux). */
void send_calc_fru_flux(double FRUdep_states[3], double *Jxfer, double *Jtr, double *ICa,
                        double *Ito2);
/* mpi_master.c:122-191 (grw_calc_avg). */
void send_calc_fru_avg(double state[37], double otherstate[12]);
/* mpi_master.c:477-485 and :493-501 (grw_save_state, grw_resume_state). */
void send_save_state(void);
void send_resume_state(void);

#ifdef __cplusplus
}
#endif
#endif
