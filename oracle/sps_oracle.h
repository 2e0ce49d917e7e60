/*
 * sps_oracle.h - CPU oracle of the Stress-Particle SPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates the reference algorithm in
 * plain C (the pair list is built in discovery order, traversed in the
 * reference's linked-list order and every pair loop scatters into both
 * endpoints, exactly as code/2_SPH_main_2018.f90 does).  It implements the
 * same C-ABI as the product (include/sps.h) so tests can drive both through
 * one harness.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.
 */
#ifndef SPS_ORACLE_H
#define SPS_ORACLE_H
#include "../include/sps.h"
#endif
