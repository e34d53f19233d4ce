// grw_mpi.cpp -- the reference's MPI-layer entry points (mpi_master.c) over one
// device context of libgrw_fru.so; see include/grw_mpi.h.  Host code only.
#include "../../include/grw_mpi.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/grw_fru.h"

// The reference program's globals (caf.c:34-36), read as mpi_master.c reads
// them.  Weak, so that the library also loads into a process that does not
// define them (a check of its exports); the calls that need them fail there.
extern "C" __attribute__((weak)) int NFRU_sim;
extern "C" __attribute__((weak)) double NFRU_scale;

namespace {

grw_ctx* g_ctx = nullptr;
int g_nfru = -1;
double g_scale = 0.0;

[[noreturn]] void die(const char* call) {
  std::fprintf(stderr, "grw_mpi: %s: %s\n", call, grw_last_error());
  std::exit(1);
}

void check(int rc, const char* call) {
  if (rc != GRW_OK) die(call);
}

grw_ctx* ctx(const char* call) {
  if (!g_ctx) {
    std::fprintf(stderr, "grw_mpi: %s: initialize_mpi_state has not been called\n", call);
    std::exit(1);
  }
  return g_ctx;
}

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

}  // namespace

extern "C" {

void initialize_mpi(int* /*argc*/, char** /*argv*/, int* rank, int* procs) {
  if (rank) *rank = 0;
  if (procs) *procs = 1;
}

void end_mpi(void) {
  if (g_ctx) (void)grw_destroy(g_ctx);
  g_ctx = nullptr;
  g_nfru = -1;
}

void parallel(void) {}

void initialize_mpi_state(double FRU_states[][5], int LType_state[][4][2], int RyR_state[][4][5],
                          int Ito2_state[][4], int mti[], unsigned long mt[][625]) {
  if (&NFRU_sim == nullptr || &NFRU_scale == nullptr) {
    std::fprintf(stderr, "grw_mpi: initialize_mpi_state: the program defines no NFRU_sim / NFRU_scale\n");
    std::exit(1);
  }
  const int n = NFRU_sim;
  if (n <= 0) {
    std::fprintf(stderr, "grw_mpi: initialize_mpi_state: NFRU_sim = %d\n", n);
    std::exit(1);
  }
  if (!g_ctx || n != g_nfru || NFRU_scale != g_scale) {
    end_mpi();
    grw_config cfg;
    grw_config_default(&cfg);
    cfg.nfru = n;
    cfg.nfru_scale = NFRU_scale;
    const char* prec = std::getenv("GRW_PRECISION");
    cfg.precision = (prec && std::strcmp(prec, "mixed") == 0) ? GRW_MIXED : GRW_FP64;
    cfg.sequential_sums = env_int("GRW_SEQUENTIAL_SUMS", 0) ? 1 : 0;
    cfg.device = env_int("GRW_DEVICE", 0);
    check(grw_create(&cfg, &g_ctx), "initialize_mpi_state");
    g_nfru = n;
    g_scale = NFRU_scale;
  }
  std::vector<uint32_t> words((size_t)n * 624);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 624; ++k) words[(size_t)i * 624 + k] = (uint32_t)mt[i][k];
  check(grw_load_state(g_ctx, &FRU_states[0][0], &LType_state[0][0][0], &RyR_state[0][0][0],
                       &Ito2_state[0][0], mti, words.data()),
        "initialize_mpi_state");
}

void parallel_get_FRUs(double FRU_states[][5], int LType_state[][4][2], int RyR_state[][4][5],
                       int Ito2_state[][4], unsigned long mt[][625], int mti[]) {
  grw_ctx* c = ctx("parallel_get_FRUs");
  std::vector<uint32_t> words((size_t)g_nfru * 624);
  check(grw_get_state(c, &FRU_states[0][0], &LType_state[0][0][0], &RyR_state[0][0][0],
                      &Ito2_state[0][0], mti, words.data()),
        "parallel_get_FRUs");
  for (int i = 0; i < g_nfru; ++i)
    for (int k = 0; k < 624; ++k) mt[i][k] = words[(size_t)i * 624 + k];
}

void distrib_simFRU(double st_time, double end_time, double FRUdep_states0[3],
                    double FRUdep_statesf[3], double* Jxfer, double* Jtr, double* ICa,
                    double* Ito2) {
  double out[4];
  check(grw_distrib_simfru(ctx("distrib_simFRU"), st_time, end_time, FRUdep_states0,
                           FRUdep_statesf, out),
        "distrib_simFRU");
  *Jxfer = out[0];
  *Jtr = out[1];
  *ICa = out[2];
  *Ito2 = out[3];
}

void send_calc_fru_flux(double FRUdep_states[3], double* Jxfer, double* Jtr, double* ICa,
                        double* Ito2) {
  double out[4];
  check(grw_calc_flux(ctx("send_calc_fru_flux"), FRUdep_states, out), "send_calc_fru_flux");
  *Jxfer = out[0];
  *Jtr = out[1];
  *ICa = out[2];
  *Ito2 = out[3];
}

void send_calc_fru_avg(double state[37], double otherstate[12]) {
  check(grw_calc_avg(ctx("send_calc_fru_avg"), state, otherstate), "send_calc_fru_avg");
}

void send_save_state(void) { check(grw_save_state(ctx("send_save_state")), "send_save_state"); }

void send_resume_state(void) {
  check(grw_resume_state(ctx("send_resume_state")), "send_resume_state");
}

}  // extern "C"
