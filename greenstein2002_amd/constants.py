"""Physical constants of the model (parameters_fcn_fru.c:21-90), for host-side
bookkeeping in Python.  The kernels use the same values from csrc/grw_model.h."""

Ko, Nao, Cao, Clo, Cli = 4.0, 138.0, 2.0, 150.0, 20.0
Faraday = 96.5
RT_over_F = 8.314 * 310.0 / 96.5
Acap = 1.534e-4
VNSR = (0.53 * 2.10e-6)
Vmyo = 25.84e-6
VJSR = (float(4) * 0.53 * 0.5 * 0.2 * 1.05e-10)
VSS = (0.5 * 0.2 * 2.03e-12)
PCa = (1.5 / 2.8 * 0.2 * 0.9 * 0.9468e-11)
PCl = 2.65e-15
BSLtot, CSQNtot, BSRtot = 1.124, 13.5, 0.047
KBSL, KmCSQN, KBSR = 0.0087, 0.63, 0.00087
LTRPNtot, HTRPNtot = 70.0e-3, 140.0e-3
CMDNtot, EGTAtot, KmCMDN, KmEGTA = 50.0e-3, 0.0, 2.38e-3, 1.5e-4
JRyRmax = (0.6 * 1.5 * 1.1 * 3.96)
tautr, tauxfer, tauss2ss = 3.0, 0.005, (10.0 * 0.005)
