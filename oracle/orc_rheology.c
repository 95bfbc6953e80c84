/*
 * orc_rheology.c -- oracle restatement of the rheology part of the MDOODZ
 * Newton-iteration path (TEST INFRASTRUCTURE: checker and CPU baseline).
 *
 *   orc_Viscosity                        RheologyDensity.c:44-707
 *   orc_RheologicalOperators             RheologyDensity.c:713-954
 *   orc_NonNewtonianViscosityGrid        RheologyDensity.c:961-1415
 *   orc_StrainRateComponents             RheologyDensity.c:1902-1982
 *   orc_InterpCentroidsToVerticesDouble  GridRoutines.c:106-196
 *   orc_EvaluateRHS                      StokesRoutines.c:922-1128
 *
 * Reference features outside this restatement (progressive reactions,
 * reac_soft; VolChangeReac) abort with a message.
 */
#include "mdz_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ZEROC 273.15
#define MINV(a, b) ((a) <= (b) ? (a) : (b))

static __thread int g_it_ve, g_it_pl;
static __thread double g_eta_up, g_eta_ve;

void orc_last_iterations(int *it_ve, int *it_pl) { *it_ve = g_it_ve; *it_pl = g_it_pl; }
void orc_last_local_solve(double *eta_up, double *eta_ve) { *eta_up = g_eta_up; *eta_ve = g_eta_ve; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* RheologyDensity.c:44 Viscosity.  Local constitutive update of one node
 * and one phase; returns the dissipative viscosity, writes the 27 outputs. */
double orc_Viscosity(int phase, double G, double T, double P, double d, double phi, double X0,
                     double Exx, double Ezz, double Exz, double Gxx, double Gzz, double Gxz,
                     double Txx0, double Tzz0, double Txz0, mat_prop *mt, params *model,
                     scale *scaling, double *txxn, double *tzzn, double *txzn, double *etaVE,
                     double *VEcoeff, double *Eii_el, double *Eii_pl, double *Eii_pwl,
                     double *Eii_exp, double *Eii_lin, double *Eii_gbs, double *Eii_cst,
                     double *Exx_el, double *Ezz_el, double *Exz_el, double *Exx_diss,
                     double *Ezz_diss, double *Exz_diss, double *d1, double strain_acc, double dil,
                     double fric, double C, double *detadexx, double *detadezz, double *detadexz,
                     double *detadp, double P0, double *X1, double *OverS, double *ddivpdexx,
                     double *ddivpdezz, double *ddivpdexz, double *ddivpdp, double *Pcorr,
                     double *drhodp, double *rho, double beta, double div, double *div_el,
                     double *div_pl, double *div_r) {
    (void)strain_acc; (void)div;
    const double R = mt->R, dt = model->dt;
    const double minEta = model->mineta, maxEta = model->maxeta;
    const double TmaxPeierls = (1200.0 + ZEROC) / scaling->T;   /* :49 */
    const int nitmax = 20;
    const double tol = 1.0e-11;
    int it;
    int constant = 0, dislocation = 0, peierls = 0, diffusion = 0, gbs = 0;
    int elastic = model->iselastic;
    double K = 1.0 / beta;
    double Tii0 = sqrt(Txx0 * Txx0 + Txz0 * Txz0);              /* :54 (xx and xz only) */

    /* flow-law parameters of this phase (:60-74) */
    double Ea_pwl = mt->Qpwl[phase], Va_pwl = mt->Vpwl[phase], n_pwl = mt->npwl[phase];
    double m_pwl = mt->mpwl[phase], r_pwl = mt->rpwl[phase], A_pwl = mt->Apwl[phase];
    double f_pwl = mt->fpwl[phase], a_pwl = mt->apwl[phase], F_pwl = mt->Fpwl[phase];
    double pre_factor = mt->pref_pwl[phase];
    double Ea_lin = mt->Qlin[phase], Va_lin = mt->Vlin[phase], n_lin = mt->nlin[phase];
    double m_lin = mt->mlin[phase], r_lin = mt->rlin[phase], A_lin = mt->Alin[phase];
    double f_lin = mt->flin[phase], a_lin = mt->alin[phase], F_lin = mt->Flin[phase];
    double Ea_gbs = mt->Qgbs[phase], Va_gbs = mt->Vgbs[phase], n_gbs = mt->ngbs[phase];
    double m_gbs = mt->mgbs[phase], r_gbs = mt->rgbs[phase], A_gbs = mt->Agbs[phase];
    double f_gbs = mt->fgbs[phase], a_gbs = mt->agbs[phase], F_gbs = mt->Fgbs[phase];
    double Ea_exp = mt->Qexp[phase], S_exp = mt->Sexp[phase], E_exp = mt->Eexp[phase];
    double gamma = mt->Gexp[phase], n_exp = mt->nexp[phase];
    int gs = mt->gs[phase];
    double pg = mt->ppzm[phase], Kg = mt->Kpzm[phase], Qg = mt->Qpzm[phase], gam = mt->Gpzm[phase];
    double cg = mt->cpzm[phase], lambda = mt->Lpzm[phase];
    double eta_vp0 = mt->eta_vp[phase], n_vp = mt->n_vp[phase], eta_vp = 0.0;
    int phase_two = mt->phase_two[phase], constant_mix = mt->phase_mix[phase];
    int mix_avg = model->diffuse_avg;
    double alpha = mt->alp[phase];

    double B_pwl = 0, B_lin = 0, B_exp = 0, B_gbs = 0, C_pwl = 0, C_lin = 0, C_exp = 0, C_gbs = 0;
    double dCpwldP = 0, dClindP = 0, ST = 0, F_exp = 0;
    double eta_pwl = 0, eta_exp = 0, eta_lin = 0, eta_el = 0, eta_gbs = 0, eta_cst = 0, eta_vep = 0;
    double eta_up, eta_lo, eta_ve, ieta_sum;
    double Txx, Tzz, Txz, Tyy, Tii, Eii, Gii, f_ani, res = 0, res0 = 0, r_eta_ve, Eii_vis, dfdeta;
    double Tyield, F_trial, gdot = 0, dQdtxx = 0, dQdtyy = 0, dQdtzz = 0, dQdtxz = 0, dQdP = 0;
    double Exx_lin = 0, Ezz_lin = 0, Exz_lin = 0, Exx_exp = 0, Ezz_exp = 0, Exz_exp = 0;
    double Exx_gbs = 0, Ezz_gbs = 0, Exz_gbs = 0, Exx_cst = 0, Ezz_cst = 0, Exz_cst = 0;
    double Exx_pl = 0, Ezz_pl = 0, Exz_pl = 0, Exx_pwl = 0, Ezz_pwl = 0, Exz_pwl = 0;
    double detadTxx, detadTzz, detadTxz, deta;
    double deta_ve_dExx, deta_ve_dEzz, deta_ve_dExz, deta_ve_dP, Jii, dummy;
    double divp = 0.0, Pc = P, F_trial0 = 0.0, dFdgdot;
    double deta_vep_dExx = 0, deta_vep_dEzz = 0, deta_vep_dExz = 0, deta_vep_dP = 0;
    double dlamdExx = 0, dlamdEzz = 0, dlamdExz = 0, dlamdP = 0;
    double ddivrdpc = 0.0, divr = 0.0, X = X0;
    int is_pl = 0;

    if (mt->reac_soft[phase] == 1) {
        fprintf(stderr, "oracle: progressive reactions (reac_soft) are outside this restatement\n");
        abort();
    }
    if (model->diffuse_X == 0) constant_mix = 0;                  /* :101 */

    /* outputs zeroed (:105-112) */
    *Eii_exp = 0; *Eii_lin = 0; *Eii_pl = 0; *Eii_pwl = 0; *Eii_el = 0; *Eii_gbs = 0; *Eii_cst = 0;
    *txxn = 0; *tzzn = 0; *txzn = 0; *etaVE = 0; *VEcoeff = 0; *Exx_el = 0; *Ezz_el = 0; *Exz_el = 0;
    *Exx_diss = 0; *Ezz_diss = 0; *Exz_diss = 0; *d1 = 0;
    *detadexx = 0; *detadezz = 0; *detadexz = 0; *detadp = 0;
    *ddivpdexx = 0; *ddivpdezz = 0; *ddivpdexz = 0; *ddivpdp = 0;
    *X1 = 0; *drhodp = 0; *rho = 0; *OverS = 0; *Pcorr = 0; *div_el = 0; *div_pl = 0; *div_r = 0;

    /* invariants (:117-123): out-of-plane component is -(xx+zz) */
    double Eyy = -(Exx + Ezz), Gyy = -(Gxx + Gzz);
    Eii = sqrt(1.0 / 2.0 * (Exx * Exx + Ezz * Ezz + Eyy * Eyy) + Exz * Exz);
    Gii = sqrt(1.0 / 2.0 * (Gxx * Gxx + Gzz * Gzz + Gyy * Gyy) + Gxz * Gxz);
    if (Eii * scaling->E < 1e-30) Eii = 1e-30 / scaling->E;
    if (Gii * scaling->E < 1e-30) Gii = 1e-30 / scaling->E;
    f_ani = Gii / Eii;
    *Pcorr = P;

    /* mechanism switches (:126-132) */
    if (mt->cstv[phase] != 0) constant = 1;
    if (mt->pwlv[phase] != 0) dislocation = 1;
    if (mt->expv[phase] != 0 && T < TmaxPeierls) peierls = 1;
    if (mt->linv[phase] != 0) diffusion = 1;
    if (mt->gbsv[phase] != 0) gbs = 1;
    if (mt->gs[phase] != 0) gs = 1;
    if (model->step == 0) elastic = 0;                             /* :135 */
    *d1 = mt->gs_ref[phase];
    if (elastic == 0) { G = 1e1; K = 1e1; dil = 0.0; }            /* :144 */
    if (T < ZEROC / scaling->T) T = ZEROC / scaling->T;           /* :147 */

    /* prefactors (:151-188) */
    if (dislocation == 1) {
        B_pwl = pre_factor * F_pwl * pow(A_pwl, -1.0 / n_pwl) * exp((Ea_pwl + P * Va_pwl) / R / n_pwl / T)
              * pow(d, m_pwl / n_pwl) * pow(f_pwl, -r_pwl / n_pwl) * exp(-a_pwl * phi / n_pwl);
        C_pwl = pow(2.0 * f_ani * B_pwl, -n_pwl);
        dCpwldP = -C_pwl * Va_pwl / R / T;
    }
    if (diffusion == 1) {
        if (m_lin > 0.0 && d < 1e-13 / scaling->L) {
            printf("Cannot run with grain size dependent viscosity if grain size is set to 0 --> d = %2.2e!!!\n", d * scaling->L);
            exit(1);
        }
        B_lin = F_lin * pow(A_lin, -1.0 / n_lin) * exp((Ea_lin + P * Va_lin) / R / n_lin / T)
              * pow(f_lin, -r_lin / n_lin) * exp(-a_lin * phi / n_lin);
        C_lin = pow(2.0 * f_ani * B_lin, -n_lin);
        dClindP = -C_lin * Va_lin / R / T;
    }
    if (gbs == 1) {
        B_gbs = F_gbs * pow(A_gbs, -1.0 / n_gbs) * exp((Ea_gbs + P * Va_gbs) / R / n_gbs / T)
              * pow(d, m_gbs / n_gbs) * pow(f_gbs, -r_gbs / n_gbs) * exp(-a_gbs * phi / n_gbs);
        C_gbs = pow(2.0 * f_ani * B_gbs, -n_gbs);
    }
    if (peierls == 1) {
        ST = Ea_exp / R / T * 2.0 * gamma * (1 - gamma);
        double Arr_exp = exp(-Ea_exp / R / T * pow(1.0 - gamma, 2.0));
        F_exp = pow(pow(2.0, 1.0 - ST - n_exp) / pow(sqrt(3.0), ST + n_exp + 1.0), 1.0 / (ST + n_exp));
        B_exp = F_exp * (pow(gamma * S_exp, ST / (ST + n_exp)) / pow(E_exp * Arr_exp, 1.0 / (ST + n_exp)));
        C_exp = pow(2.0 * f_ani * B_exp, -(ST + n_exp));
    }

    /* Drucker-Prager yield, von-Mises and tension cut-offs (:190-214) */
    double cos_fric = cos(fric), sin_fric = sin(fric), sin_dil = sin(dil);
    Tyield = C * cos_fric + P * sin_fric;
    if (mt->Slim[phase] < Tyield) {
        sin_fric = 0.0; cos_fric = 1.0; sin_dil = 0.0;
        C = mt->Slim[phase]; Tyield = mt->Slim[phase];
    }
    int is_tensile = mt->is_tensile[phase], tens = 0;
    double P_tens = -25e6 / scaling->S;
    double sin_fric_tens = -C * cos_fric / P_tens;
    if (C * cos_fric + P * sin_fric_tens < Tyield && is_tensile == 1) {
        sin_fric = sin_fric_tens;
        Tyield = C * cos_fric + P * sin_fric_tens;
        tens = 1;
    }

    X = X0;
    *X1 = X;

    /* isolated viscosities (:299-308) */
    eta_el = G * dt;
    if (constant == 1) eta_cst = mt->eta0[phase];
    if (constant_mix == 1 && mix_avg == 0) eta_cst = X0 * mt->eta0[phase] + (1 - X0) * mt->eta0[phase_two];
    if (constant_mix == 1 && mix_avg == 1) eta_cst = pow(X0 / mt->eta0[phase] + (1 - X0) / mt->eta0[phase_two], -1.0);
    if (constant_mix == 1 && mix_avg == 2) eta_cst = exp(X0 * log(mt->eta0[phase]) + (1 - X0) * log(mt->eta0[phase_two]));
    if (dislocation == 1) eta_pwl = B_pwl * pow(Eii, 1.0 / n_pwl - 1.0);
    if (diffusion == 1) eta_lin = B_lin * pow(Eii, 1.0 / n_lin - 1.0) * pow(d, m_lin / n_lin);
    if (gbs == 1) eta_gbs = B_gbs * pow(Eii, 1.0 / n_gbs - 1.0);
    if (peierls == 1) eta_exp = B_exp * pow(Eii, 1.0 / (ST + n_exp) - 1.0);

    /* bounds and initial guess (:318-355): eta_up = min of isolated ones */
    *Eii_pl = 0.0;
    eta_up = 1.0e100 / scaling->eta;
    ieta_sum = 0.0;
    if (constant == 1) { eta_up = MINV(eta_up, eta_cst); ieta_sum += 1.0 / eta_cst; }
    if (dislocation == 1) { eta_up = MINV(eta_up, eta_pwl); ieta_sum += 1.0 / eta_pwl; }
    if (elastic == 1) { eta_up = MINV(eta_up, eta_el); ieta_sum += 1.0 / eta_el; }
    if (peierls == 1) { eta_up = MINV(eta_up, eta_exp); ieta_sum += 1.0 / eta_exp; }
    if (diffusion == 1) { eta_up = MINV(eta_up, eta_lin); ieta_sum += 1.0 / eta_lin; }
    if (gbs == 1) { eta_up = MINV(eta_up, eta_gbs); ieta_sum += 1.0 / eta_gbs; }
    eta_lo = 1.0 / ieta_sum;
    (void)eta_lo;
    eta_ve = eta_up;

    /* visco-elastic local Newton iterations on eta_ve (:358-387) */
    for (it = 0; it < nitmax; it++) {
        Tii = 2.0 * f_ani * eta_ve * Eii;
        if (constant == 1) *Eii_cst = Tii / f_ani / 2.0 / eta_cst;
        if (dislocation == 1) *Eii_pwl = C_pwl * pow(Tii, n_pwl);
        if (gbs == 1) *Eii_gbs = C_gbs * pow(Tii, n_gbs);
        if (peierls == 1) *Eii_exp = C_exp * pow(Tii, ST + n_exp);
        if (gs == 1) *d1 = exp(log(Kg * exp(-Qg / R / T) * gam / (lambda * (1.0 / cg) * Tii * (*Eii_pwl + *Eii_exp + *Eii_gbs + *Eii_pl) * pg)) / (1.0 + pg));
        if (diffusion == 1) *Eii_lin = C_lin * pow(Tii, n_lin) * pow(*d1, -m_lin);
        Eii_vis = *Eii_pwl + *Eii_exp + *Eii_lin + *Eii_gbs + *Eii_cst;
        r_eta_ve = Eii - elastic * Tii / (2.0 * eta_el) - Eii_vis;
        res = fabs(r_eta_ve / Eii);
        if (it == 0) res0 = res;
        if (res < tol / 100) break;
        dfdeta = 0.0;
        if (elastic == 1) dfdeta += -Eii / eta_el;
        if (peierls == 1) dfdeta += -(*Eii_exp) * (ST + n_exp) / eta_ve;
        if (diffusion == 1) dfdeta += -(*Eii_lin) * n_lin / eta_ve;
        if (dislocation == 1) dfdeta += -(*Eii_pwl) * n_pwl / eta_ve;
        if (constant == 1) dfdeta += -Eii / eta_cst;
        eta_ve -= r_eta_ve / dfdeta;
    }
    g_it_ve = it;
    g_eta_up = eta_up;
    g_eta_ve = eta_ve;
    (void)res0;
    /* :389 -- the check can only trigger on it == nitmax-1, as in the reference */
    if (it == nitmax - 1 && res > tol) { printf("Visco-Elastic iterations failed!\n"); exit(0); }

    /* stresses (:393-397) */
    Tii = 2.0 * eta_ve * f_ani * Eii;
    Txx = Gxx * Tii / f_ani / Eii;
    Tyy = Gxx * Tii / f_ani / Eii;
    Tzz = Gzz * Tii / f_ani / Eii;
    Txz = Gxz * Tii / f_ani / Eii;

    if (dislocation == 1) eta_pwl = pow(2.0 * C_pwl, -1.0) * pow(Tii, 1.0 - n_pwl);
    if (diffusion == 1) eta_lin = B_lin * pow(*Eii_lin, 1.0 / n_lin - 1.0) * pow(*d1, m_lin / n_lin);
    if (peierls == 1) eta_exp = B_exp * pow(*Eii_exp, 1.0 / (ST + n_exp) - 1.0);
    if (elastic == 1) *Exx_el = (double)elastic * (Txx - Txx0) / 2.0 / eta_el;
    if (elastic == 1) *Ezz_el = (double)elastic * (Tzz - Tzz0) / 2.0 / eta_el;
    if (elastic == 1) *Exz_el = (double)elastic * (Txz - Txz0) / 2.0 / eta_el;
    if (dislocation == 1) { Exx_pwl = Txx / 2.0 / eta_pwl; Ezz_pwl = Tzz / 2.0 / eta_pwl; Exz_pwl = Txz / 2.0 / eta_pwl; }
    if (diffusion == 1) { Exx_lin = Txx / 2.0 / eta_lin; Ezz_lin = Tzz / 2.0 / eta_lin; Exz_lin = Txz / 2.0 / eta_lin; }
    if (peierls == 1) { Exx_exp = Txx / 2.0 / eta_exp; Ezz_exp = Tzz / 2.0 / eta_exp; Exz_exp = Txz / 2.0 / eta_exp; }
    if (constant == 1) { Exx_cst = Txx / 2.0 / eta_cst; Ezz_cst = Tzz / 2.0 / eta_cst; Exz_cst = Txz / 2.0 / eta_cst; }

    /* visco-elastic consistent tangent (:425-447) */
    deta = 0.0;
    if (dislocation == 1) deta += -C_pwl * pow(Tii, n_pwl - 3.0) * (n_pwl - 1.0);
    if (peierls == 1) deta += -C_exp * pow(Tii, n_exp + ST - 3.0) * (n_exp + ST - 1.0);
    if (diffusion == 1) deta += -C_lin * pow(Tii, n_lin - 3.0) * (n_lin - 1.0) * pow(*d1, -m_lin);
    detadTxx = deta * (2.0 * Txx + Tzz) * pow(eta_ve, 2.0);
    detadTzz = deta * (2.0 * Tzz + Txx) * pow(eta_ve, 2.0);
    detadTxz = 2.0 * deta * Txz * pow(eta_ve, 2.0);
    double den = 1.0 - 2.0 * (detadTxx * Exx + detadTzz * Ezz + detadTxz * Exz);
    deta_ve_dExx = detadTxx * 2.0 * eta_ve / den;
    deta_ve_dEzz = detadTzz * 2.0 * eta_ve / den;
    deta_ve_dExz = detadTxz * 2.0 * eta_ve / den;
    dummy = 0.0;
    Jii = Tii * Tii;
    if (dislocation == 1) dummy += -2.0 * pow(Jii, (1.0 / 2.0) * n_pwl - 1.0 / 2.0) * dCpwldP;
    if (diffusion == 1) dummy += -2.0 * pow(Jii, (1.0 / 2.0) * n_lin - 1.0 / 2.0) * pow(d, -m_lin) * dClindP;
    deta_ve_dP = pow(eta_ve, 2) * dummy;

    /* yield check and return map in the plastic multiplier (:468-531) */
    F_trial = Tii - Tyield;
    double Pc_chk, F_corr1;
    double sin_dil_t = sin_dil;
    if (tens == 1) {
        eta_vp = eta_vp0 * pow(Eii, 1.0 / n_vp - 1);
        Pc_chk = -(C * cos_fric) / (-Tii / P + sin_fric);
        sin_dil_t = (P * eta_ve + P * eta_vp - Pc_chk * eta_ve - Pc_chk * eta_vp) / (K * dt * (C * cos_fric + Pc_chk * sin_fric - Tii));
        sin_dil = sin_dil_t;
    }
    g_it_pl = 0;
    if (F_trial > 1e-17) {
        is_pl = 1;
        eta_vp = eta_vp0 * pow(Eii, 1.0 / n_vp - 1);
        gdot = F_trial / (eta_ve + eta_vp + K * dt * sin_fric * sin_dil);
        dQdtxx = Txx / 2.0 / Tii;
        dQdtyy = Tyy / 2.0 / Tii;
        dQdtzz = Tzz / 2.0 / Tii;
        dQdtxz = Txz / 1.0 / Tii;
        dQdP = -sin_dil;
        res0 = F_trial;
        F_trial0 = F_trial;
        for (it = 0; it < nitmax; it++) {
            dQdP = -sin_dil;
            divp = -gdot * dQdP;
            Pc = P + K * dt * divp;
            eta_vp = eta_vp0 * pow(fabs(gdot), 1.0 / n_vp - 1.0);
            Tyield = C * cos_fric + Pc * sin_fric + gdot * eta_vp;
            F_trial = Tii - eta_ve * gdot - Tyield;
            res = fabs(F_trial);
            if (res < tol || res / F_trial0 < tol) break;
            dFdgdot = -eta_ve - eta_vp / n_vp - K * dt * sin_fric * sin_dil;
            gdot -= F_trial / dFdgdot;
        }
        g_it_pl = it + 1;
        if (it == nitmax - 1 && (res > tol || res / F_trial0 > tol)) { printf("Visco-Plastic iterations failed!\n"); exit(0); }
        Txx = 2.0 * eta_ve * (Exx - gdot * dQdtxx);
        Tzz = 2.0 * eta_ve * (Ezz - gdot * dQdtzz);
        Tyy = -(Txx + Tzz);
        Txz = 2.0 * eta_ve * (Exz - gdot * dQdtxz / 2.0);
        Tii = sqrt(0.5 * (pow(Txx, 2.0) + pow(Tzz, 2.0) + pow(Tyy, 2.0)) + pow(Txz, 2.0));
        eta_vep = Tii / (2.0 * Eii);
        F_corr1 = Tii - Tyield;
        *Eii_pl = gdot / 2.0;
        Tii = 2.0 * eta_vep * Eii;
        *OverS = eta_vp * gdot;
        *Pcorr = Pc;
        /* RheologyDensity.c:523-530 consistency warning (printed only when
         * MDZ_ORACLE_VERBOSE is set; the check itself changes nothing) */
        if (is_pl == 1 && (fabs(F_corr1) > 1e-7) && (fabs(F_corr1) / res0 > 1e-7) && getenv("MDZ_ORACLE_VERBOSE")) {
            printf("Problem with F_corr in phase = %d\n", phase);
            printf("it=%d tens=%d Tyield=%g F_trial0=%g res=%g F_corr1=%g Tii=%g Pc=%g P=%g sin_fric=%g sin_dil=%g K=%g gdot=%g eta_ve=%g\n",
                   it, tens, Tyield, res0, res, F_corr1, Tii, Pc, P, sin_fric, sin_dil, K, gdot, eta_ve);
            printf("WARNING: Non linear convergence is expected to be poor\n");
        }
    }
    (void)dQdtyy;

    /* visco-elasto-plastic consistent tangent (:538-572) */
    double dsin_dildExx = 0.0, dsin_dildEzz = 0.0, dsin_dildExz = 0.0, dsin_dildP = 0.0;
    if (is_pl == 1) {
        if (tens == 1) {
            double dEii_dExx = (2.0 * Exx + Ezz) / 2.0 / Eii;
            double dEii_dEzz = (2.0 * Ezz + Exx) / 2.0 / Eii;
            double dEii_dExz = Exz / 1.0 / Eii;
            /* derivative of the tensile dilation slope; same closed form for
             * the three strain-rate directions with (deta_ve_dE, dEii_dE) */
            double a2 = 2 * Eii * eta_ve - P * sin_fric;
            double bden = K * dt * a2 * (C * P * cos_fric * sin_fric + (C * cos_fric - 2 * Eii * eta_ve) * a2);
            double q1 = C * P * cos_fric * sin_fric + a2 * a2;
            double de[3] = {deta_ve_dExx, deta_ve_dEzz, deta_ve_dExz};
            double dE[3] = {dEii_dExx, dEii_dEzz, dEii_dExz};
            double outd[3];
            for (int i = 0; i < 3; i++) {
                double dT = Eii * de[i] + dE[i] * eta_ve;
                outd[i] = (2 * K * dt * sin_dil * dT * q1
                           + P * (-C * cos_fric * de[i] * a2 + 2 * C * cos_fric * (eta_ve + eta_vp) * dT + de[i] * a2 * a2)) / bden;
            }
            dsin_dildExx = outd[0]; dsin_dildEzz = outd[1]; dsin_dildExz = outd[2];
            dsin_dildP = (2 * C * Eii * cos_fric * (eta_ve + eta_vp) * (P * deta_ve_dP - eta_ve)
                          - C * P * cos_fric * deta_ve_dP * a2
                          + 2 * Eii * K * dt * sin_dil * (C * cos_fric * sin_fric * (P * deta_ve_dP - eta_ve) + deta_ve_dP * a2 * a2)
                          + a2 * a2 * (P * deta_ve_dP + eta_ve + eta_vp)) / bden;
        }
        double dFdExx = (2.0 * Exx + Ezz) * eta_ve / Eii + 2.0 * Eii * deta_ve_dExx;
        double dFdEzz = (2.0 * Ezz + Exx) * eta_ve / Eii + 2.0 * Eii * deta_ve_dEzz;
        double dFdExz = 2.0 * Exz * eta_ve / Eii + 2.0 * Eii * deta_ve_dExz;
        double dFdP = -sin_fric + 2.0 * Eii * deta_ve_dP;
        double g = 1.0 / (eta_ve + eta_vp / n_vp + K * dt * sin_fric * sin_dil);
        dlamdExx = g * (dFdExx - gdot * (deta_ve_dExx + K * dt * sin_fric * dsin_dildExx));
        dlamdEzz = g * (dFdEzz - gdot * (deta_ve_dEzz + K * dt * sin_fric * dsin_dildEzz));
        dlamdExz = g * (dFdExz - gdot * (deta_ve_dExz + K * dt * sin_fric * dsin_dildExz));
        dlamdP = g * (dFdP - gdot * (deta_ve_dP + K * dt * sin_fric * dsin_dildP));
        double deta_vp_dExx = eta_vp / gdot * dlamdExx * (1.0 / n_vp - 1.0);
        double deta_vp_dEzz = eta_vp / gdot * dlamdEzz * (1.0 / n_vp - 1.0);
        double deta_vp_dExz = eta_vp / gdot * dlamdExz * (1.0 / n_vp - 1.0);
        double deta_vp_dP = eta_vp / gdot * dlamdP * (1.0 / n_vp - 1.0);
        double a = eta_vp + K * dt * sin_fric * sin_dil;
        deta_vep_dExx = -0.5 * (2.0 * Exx + Ezz) * Tyield / (2.0 * pow(Eii, 3.0)) + (a * dlamdExx + gdot * deta_vp_dExx + gdot * K * dt * sin_fric * dsin_dildExx) / (2.0 * Eii);
        deta_vep_dEzz = -0.5 * (2.0 * Ezz + Exx) * Tyield / (2.0 * pow(Eii, 3.0)) + (a * dlamdEzz + gdot * deta_vp_dEzz + gdot * K * dt * sin_fric * dsin_dildEzz) / (2.0 * Eii);
        deta_vep_dExz = -Exz * Tyield / (2.0 * pow(Eii, 3.0)) + (a * dlamdExz + gdot * deta_vp_dExz + gdot * K * dt * sin_fric * dsin_dildExz) / (2.0 * Eii);
        deta_vep_dP = (sin_fric + a * dlamdP + gdot * deta_vp_dP + gdot * K * dt * sin_fric * dsin_dildP) / (2.0 * Eii);
    }

    /* density (:576-593, non-reactive branch) */
    {
        double rho_ref = mt->rho[phase];
        *rho = rho_ref * exp(1.0 / K * Pc - alpha * T);
        *drhodp = (*rho) / K;
    }

    /* outputs (:605-699) */
    double inv_eta_diss = 0.0;
    if (peierls == 1) inv_eta_diss += (1.0 / eta_exp);
    if (dislocation == 1) inv_eta_diss += (1.0 / eta_pwl);
    if (diffusion == 1) inv_eta_diss += (1.0 / eta_lin);
    if (constant == 1) inv_eta_diss += (1.0 / eta_cst);
    if (is_pl == 0) {
        *detadexx = deta_ve_dExx; *detadezz = deta_ve_dEzz; *detadexz = deta_ve_dExz; *detadp = deta_ve_dP;
        *etaVE = eta_ve;
        *ddivpdexx = 0.0; *ddivpdezz = 0.0; *ddivpdexz = 0.0; *ddivpdp = ddivrdpc;
    } else {
        *detadexx = deta_vep_dExx; *detadezz = deta_vep_dEzz; *detadexz = deta_vep_dExz; *detadp = deta_vep_dP;
        *etaVE = eta_vep;
        *ddivpdexx = -dQdP * dlamdExx + gdot * dsin_dildExx;
        *ddivpdezz = -dQdP * dlamdEzz + gdot * dsin_dildEzz;
        *ddivpdexz = -dQdP * dlamdExz + gdot * dsin_dildExz;
        *ddivpdp = -dQdP * dlamdP + gdot * dsin_dildP + ddivrdpc;
        inv_eta_diss += 1.0 / eta_vep;
    }
    /* unconditional power-law strain-rate partition (:643-653); with no
     * dislocation creep C_pwl = 0 and eta_pwl = +inf, the partition is 0 */
    eta_pwl = pow(2.0 * C_pwl, -1.0) * pow(Tii, 1.0 - n_pwl);
    *Exx_el = (double)elastic * (Txx - Txx0) / 2.0 / eta_el;
    *Ezz_el = (double)elastic * (Tzz - Tzz0) / 2.0 / eta_el;
    *Exz_el = (double)elastic * (Txz - Txz0) / 2.0 / eta_el;
    Exx_pl = gdot * dQdtxx;
    Ezz_pl = gdot * dQdtzz;
    Exz_pl = gdot * dQdtxz / 2.0;
    Exx_pwl = Txx / 2.0 / eta_pwl;
    Ezz_pwl = Tzz / 2.0 / eta_pwl;
    Exz_pwl = Txz / 2.0 / eta_pwl;
    *Exx_diss = Exx_pl + Exx_lin + Exx_pwl + Exx_exp + Exx_gbs + Exx_cst;
    *Ezz_diss = Ezz_pl + Ezz_lin + Ezz_pwl + Ezz_exp + Ezz_gbs + Ezz_cst;
    *Exz_diss = Exz_pl + Exz_lin + Exz_pwl + Exz_exp + Exz_gbs + Exz_cst;
    *Eii_el = (double)elastic * fabs(Tii - Tii0) / 2.0 / eta_el;
    *Eii_pwl = Tii / 2.0 / eta_pwl;
    *div_pl = divp;
    *div_el = -(Pc - P0) / (K * dt);
    *div_r = divr;
    double eta = 1.0 / (inv_eta_diss);
    *VEcoeff = eta_ve / eta_el;
    if (elastic == 0) *VEcoeff = 0.0;
    if (model->step == 0) *etaVE = eta;
    *txxn = Txx; *tzzn = Tzz; *txzn = Txz;
    if (eta > maxEta) eta = maxEta;
    if (eta < minEta) eta = minEta;
    if (*etaVE > maxEta) *etaVE = maxEta;
    if (*etaVE < minEta) *etaVE = minEta;
    return eta;
}

/* GridRoutines.c:106 InterpCentroidsToVerticesDouble: pad the centroid
 * field by one ring (zero-flux copies, or periodic in x), four-point mean
 * to every non-air vertex. */
void orc_InterpCentroidsToVerticesDouble(double *cen, double *ver, grid *mesh, params *model) {
    int per = model->isperiodic_x;
    int Nx = mesh->Nx, Nz = mesh->Nz, Ncx = Nx - 1, Ncz = Nz - 1, W = Ncx + 2;
    double *t = (double *)calloc((size_t)(Ncx + 2) * (Ncz + 2), sizeof(double));
    for (int k = 0; k < Ncx; k++)
        for (int l = 0; l < Ncz; l++) t[k + (l + 1) * W + 1] = cen[k + l * Ncx];
    for (int k = 1; k < Ncx + 1; k++) t[k] = t[k + W];
    for (int k = 1; k < Ncx + 1; k++) t[k + (Ncz + 1) * W] = t[k + Ncz * W];
    for (int l = 1; l < Ncz + 1; l++) t[l * W] = per ? t[(Ncx + 2 - 1 - 1) + l * W] : t[1 + l * W];
    for (int l = 1; l < Ncz + 1; l++) t[(Ncx + 1) + l * W] = per ? t[1 + l * W] : t[Ncx + l * W];
    t[0] = per ? t[W] : t[1 + W];
    t[Ncx + 1] = per ? t[(Ncx + 1) + W] : t[Ncx + W];
    t[(Ncz + 1) * W] = per ? t[Ncz * W] : t[1 + Ncz * W];
    t[(Ncx + 1) + (Ncz + 1) * W] = per ? t[(Ncx + 1) + Ncz * W] : t[Ncx + Ncz * W];
#pragma omp parallel for schedule(static)
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx; k++) {
            int c1 = k + l * Nx, c0 = k + (l + 1) * W + 1;
            ver[c1] = 0.0;
            if (mesh->BCg.type[c1] != 30)
                ver[c1] = 0.25 * (t[c0 - 1 - W] + t[c0 - 0 - W] + t[c0 - 1] + t[c0]);
        }
    free(t);
}

/* RheologyDensity.c:1902 StrainRateComponents */
void orc_StrainRateComponents(grid *mesh, scale scaling, params *model) {
    (void)scaling;
    int Nx = mesh->Nx, Nz = mesh->Nz, Ncx = Nx - 1, Ncz = Nz - 1;
    double dx = mesh->dx, dz = mesh->dz;
#pragma omp parallel for schedule(static)
    for (int c0 = 0; c0 < Ncx * Ncz; c0++) {
        int k = c0 % Ncx, l = c0 / Ncx;
        int c1 = k + l * Nx, c2 = k + l * (Nx + 1);
        if (mesh->BCp.type[c0] != 30 && mesh->BCp.type[c0] != 31) {
            mesh->div_u[c0] = (mesh->u_in[c1 + 1 + Nx] - mesh->u_in[c1 + Nx]) / dx + (mesh->v_in[c2 + Nx + 1 + 1] - mesh->v_in[c2 + 1]) / dz;
            mesh->exxd[c0] = (mesh->u_in[c1 + 1 + Nx] - mesh->u_in[c1 + Nx]) / dx - 1.0 / 3.0 * mesh->div_u[c0];
            mesh->ezzd[c0] = (mesh->v_in[c2 + 1 + (Nx + 1)] - mesh->v_in[c2 + 1]) / dz - 1.0 / 3.0 * mesh->div_u[c0];
        } else {
            mesh->div_u[c0] = 0.0; mesh->exxd[c0] = 0.0; mesh->ezzd[c0] = 0.0;
        }
    }
#pragma omp parallel for schedule(static)
    for (int c1 = 0; c1 < Nx * Nz; c1++) {
        int k = c1 % Nx, l = c1 / Nx;
        int c2 = k + l * (Nx + 1);
        mesh->exz[c1] = 0.0;
        if (mesh->BCg.type[c1] != 30) {
            if (mesh->BCu.type[c1] != 30 && mesh->BCu.type[c1 + Nx] != 30) mesh->exz[c1] += 0.5 * (mesh->u_in[c1 + Nx] - mesh->u_in[c1]) / dz;
            if (mesh->BCv.type[c2] != 30 && mesh->BCv.type[c2 + 1] != 30) mesh->exz[c1] += 0.5 * (mesh->v_in[c2 + 1] - mesh->v_in[c2]) / dx;
        }
    }
#pragma omp parallel for schedule(static)
    for (int c0 = 0; c0 < Ncx * Ncz; c0++) {
        int k = c0 % Ncx, l = c0 / Ncx;
        int c1 = k + l * Nx;
        double sum = 0.0;
        mesh->exz_n[c0] = 0.0;
        if (mesh->BCp.type[c0] != 30 && mesh->BCp.type[c0] != 31) {
            if (mesh->BCg.type[c1] != 30) { mesh->exz_n[c0] += mesh->exz[c1]; sum++; }
            if (mesh->BCg.type[c1 + 1] != 30) { mesh->exz_n[c0] += mesh->exz[c1 + 1]; sum++; }
            if (mesh->BCg.type[c1 + Nx] != 30) { mesh->exz_n[c0] += mesh->exz[c1 + Nx]; sum++; }
            if (mesh->BCg.type[c1 + Nx + 1] != 30) { mesh->exz_n[c0] += mesh->exz[c1 + Nx + 1]; sum++; }
            if (sum > 0) mesh->exz_n[c0] /= sum;
        }
    }
    orc_InterpCentroidsToVerticesDouble(mesh->exxd, mesh->exxd_s, mesh, model);
    orc_InterpCentroidsToVerticesDouble(mesh->ezzd, mesh->ezzd_s, mesh, model);
}

/* Effective visco-elastic strain rates of one node through the inverse of
 * the anisotropy tensor (RheologyDensity.c:1061-1094); isotropic case gives
 * E = edot + tau0 / (2 G dt). */
static void ve_strain_rates(int aniso, int aniso_fstrain, double nx, double nz, double af, double fs,
                            double exx, double ezz, double exz, double sxx0, double szz0, double sxz0,
                            double etae, double *Exx, double *Ezz, double *Exz, double *Gxx,
                            double *Gzz, double *Gxz) {
    double ani = 0.0, d0 = 0.0, d1 = 0.0;
    if (aniso != 0) {
        if (aniso_fstrain == 0) ani = 1.0 - 1.0 / af;
        if (aniso_fstrain == 1) ani = 1.0 - 1.0 / fs;
        d0 = 2.0 * pow(nx, 2.0) * pow(nz, 2.0);
        d1 = nx * nz * (-pow(nx, 2.0) + pow(nz, 2.0));
    }
    double Da11 = 2.0 - 2.0 * ani * d0, Da12 = 2.0 * ani * d0, Da13 = 2.0 * ani * d1;
    double Da22 = 2.0 - 2.0 * ani * d0, Da23 = -2.0 * ani * d1, Da33 = 1.0 + 2.0 * ani * (d0 - 0.5);
    double a11 = Da33 * Da22 - pow(Da23, 2), a12 = Da13 * Da23 - Da33 * Da12, a13 = Da12 * Da23 - Da13 * Da22;
    double a22 = Da33 * Da11 - pow(Da13, 2), a23 = Da12 * Da13 - Da11 * Da23, a33 = Da11 * Da22 - pow(Da12, 2);
    double det = (Da11 * a11) + (Da12 * a12) + (Da13 * a13);
    double i11 = a11 / det, i12 = a12 / det, i13 = a13 / det, i22 = a22 / det, i23 = a23 / det, i33 = a33 / det;
    *Exx = exx + (i11 * sxx0 + i12 * szz0 + i13 * sxz0) / etae;
    *Ezz = ezz + (i12 * sxx0 + i22 * szz0 + i23 * sxz0) / etae;
    *Exz = exz + (i13 * sxx0 + i23 * szz0 + i33 * sxz0) / 2.0 / etae;
    double gxz = 2.0 * exz + (i13 * sxx0 + i23 * szz0 + i33 * sxz0) / etae;
    *Gxx = *Exx * (1.0 - ani * d0) + *Ezz * ani * d0 + gxz * ani * d1;
    *Gzz = *Ezz * (1.0 - ani * d0) + *Exx * ani * d0 - gxz * ani * d1;
    *Gxz = *Exx * ani * d1 - *Exx * ani * d1 + gxz * (ani * (d0 - 0.5) + 0.5);
}

/* RheologyDensity.c:961 NonNewtonianViscosityGrid */
void orc_NonNewtonianViscosityGrid(grid *mesh, mat_prop *materials, params *model, Nparams Nmodel,
                                   scale *scaling) {
    (void)Nmodel;
    int Nx = mesh->Nx, Nz = mesh->Nz, Ncx = Nx - 1, Ncz = Nz - 1;
    int average = model->eta_avg, UnsplitDiffReac = model->UnsplitDiffReac;
    if (model->VolChangeReac == 1) {
        fprintf(stderr, "oracle: VolChangeReac is outside this restatement\n");
        abort();
    }
    orc_InterpCentroidsToVerticesDouble(mesh->div_u, mesh->div_u_s, mesh, model);
    orc_InterpCentroidsToVerticesDouble(mesh->T, mesh->T_s, mesh, model);
    orc_InterpCentroidsToVerticesDouble(mesh->p_in, mesh->P_s, mesh, model);
    orc_InterpCentroidsToVerticesDouble(mesh->d0_n, mesh->d0_s, mesh, model);
    orc_InterpCentroidsToVerticesDouble(mesh->phi0_n, mesh->phi0_s, mesh, model);

#pragma omp parallel for schedule(dynamic, 256)
    for (int c0 = 0; c0 < Ncx * Ncz; c0++) {
        double eta = 0, txx1 = 0, tzz1 = 0, txz1 = 0, etaVE = 0, VEcoeff = 0, eII_el = 0, eII_pl = 0;
        double eII_pwl = 0, eII_exp = 0, eII_lin = 0, eII_gbs = 0, eII_cst = 0, dnew = 0;
        double exx_el = 0, ezz_el = 0, exz_el = 0, exx_diss = 0, ezz_diss = 0, exz_diss = 0;
        double detadexx = 0, detadezz = 0, detadexz = 0, detadp = 0, Xreac = 0, OverS = 0;
        double ddivpdexx = 0, ddivpdezz = 0, ddivpdexz = 0, ddivpdp = 0, Pcorr = 0, drhodp = 0, rho = 0;
        double div_el = 0, div_pl = 0, div_r = 0;
        mesh->eta_n[c0] = 0.0; mesh->eta_phys_n[c0] = 0.0; mesh->VE_n[c0] = 0.0;
        mesh->sxxd[c0] = 0.0; mesh->szzd[c0] = 0.0; mesh->sxz_n[c0] = 0.0;
        mesh->eII_el[c0] = 0.0; mesh->eII_pl[c0] = 0.0; mesh->eII_pwl[c0] = 0.0; mesh->eII_exp[c0] = 0.0;
        mesh->eII_lin[c0] = 0.0; mesh->eII_gbs[c0] = 0.0; mesh->eII_cst[c0] = 0.0; mesh->d_n[c0] = 0.0;
        mesh->exx_el[c0] = 0.0; mesh->exx_diss[c0] = 0.0; mesh->ezz_el[c0] = 0.0; mesh->ezz_diss[c0] = 0.0;
        mesh->detadexx_n[c0] = 0.0; mesh->detadezz_n[c0] = 0.0; mesh->detadgxz_n[c0] = 0.0; mesh->detadp_n[c0] = 0.0;
        mesh->ddivpdexx_n[c0] = 0.0; mesh->ddivpdezz_n[c0] = 0.0; mesh->ddivpdgxz_n[c0] = 0.0; mesh->ddivpdp_n[c0] = 0.0;
        mesh->p_corr[c0] = 0.0; mesh->div_u_el[c0] = 0.0; mesh->div_u_pl[c0] = 0.0; mesh->div_u_r[c0] = 0.0;
        mesh->Wtot[c0] = 0.0; mesh->Wel[c0] = 0.0; mesh->Wdiss[c0] = 0.0;
        if (UnsplitDiffReac == 0) mesh->X_n[c0] = 0.0;
        mesh->OverS_n[c0] = 0.0;
        if (mesh->BCp.type[c0] == 30 || mesh->BCp.type[c0] == 31) continue;

        double etae = model->iselastic == 1 ? model->dt * mesh->mu_n[c0] : 1.0;
        double Exx, Ezz, Exz, Gxx, Gzz, Gxz;
        ve_strain_rates(model->aniso, model->aniso_fstrain, mesh->nx0_n[c0], mesh->nz0_n[c0],
                        mesh->aniso_factor_n[c0], mesh->FS_AR_n[c0], mesh->exxd[c0], mesh->ezzd[c0],
                        mesh->exz_n[c0], mesh->sxxd0[c0], mesh->szzd0[c0], mesh->sxz0_n[c0], etae,
                        &Exx, &Ezz, &Exz, &Gxx, &Gzz, &Gxz);
        for (int p = 0; p < model->Nb_phases; p++) {
            double fr = mesh->phase_perc_n[p][c0];
            int cond = fabs(fr) > 1.0e-13;
            if (!cond) continue;
            eta = orc_Viscosity(p, mesh->mu_n[c0], mesh->T[c0], mesh->p_in[c0], mesh->d0_n[c0], mesh->phi0_n[c0], mesh->X0_n[c0],
                                Exx, Ezz, Exz, Gxx, Gzz, Gxz, mesh->sxxd0[c0], mesh->szzd0[c0], mesh->sxz0_n[c0],
                                materials, model, scaling, &txx1, &tzz1, &txz1, &etaVE, &VEcoeff, &eII_el, &eII_pl,
                                &eII_pwl, &eII_exp, &eII_lin, &eII_gbs, &eII_cst, &exx_el, &ezz_el, &exz_el,
                                &exx_diss, &ezz_diss, &exz_diss, &dnew, mesh->strain_n[c0], mesh->dil_n[c0],
                                mesh->fric_n[c0], mesh->C_n[c0], &detadexx, &detadezz, &detadexz, &detadp,
                                mesh->p0_n[c0], &Xreac, &OverS, &ddivpdexx, &ddivpdezz, &ddivpdexz, &ddivpdp,
                                &Pcorr, &drhodp, &rho, mesh->bet_n[c0], mesh->div_u[c0], &div_el, &div_pl, &div_r);
            if (average == 0) {
                mesh->eta_n[c0] += fr * etaVE;
                mesh->eta_phys_n[c0] += fr * eta;
                mesh->detadexx_n[c0] += fr * detadexx;
                mesh->detadezz_n[c0] += fr * detadezz;
                mesh->detadgxz_n[c0] += fr * detadexz / 2.0;
                mesh->detadp_n[c0] += fr * detadp;
            }
            if (average == 0 || average == 2) {
                mesh->sxxd[c0] += fr * txx1;
                mesh->szzd[c0] += fr * tzz1;
                mesh->sxz_n[c0] += fr * txz1;
            }
            mesh->VE_n[c0] += fr * VEcoeff;
            mesh->eII_el[c0] += fr * eII_el;
            mesh->eII_pl[c0] += fr * eII_pl;
            mesh->eII_pwl[c0] += fr * eII_pwl;
            mesh->eII_exp[c0] += fr * eII_exp;
            mesh->eII_lin[c0] += fr * eII_lin;
            mesh->eII_gbs[c0] += fr * eII_gbs;
            mesh->eII_cst[c0] += fr * eII_cst;
            mesh->d_n[c0] += fr * 1.0 / dnew;
            mesh->exx_el[c0] += fr * exx_el;
            mesh->exx_diss[c0] += fr * exx_diss;
            mesh->ezz_el[c0] += fr * ezz_el;
            mesh->ezz_diss[c0] += fr * ezz_diss;
            double exxd = mesh->exxd[c0], ezzd = mesh->ezzd[c0];
            mesh->Wtot[c0] += fr * (txx1 * exxd + tzz1 * ezzd + (txx1 + tzz1) * (exxd + ezzd) + 2.0 * txz1 * mesh->exz_n[c0]);
            mesh->Wdiss[c0] += fr * (txx1 * exx_diss + tzz1 * ezz_diss + (txx1 + tzz1) * (exx_diss + ezz_diss) + 2.0 * txz1 * exz_diss);
            mesh->Wel[c0] += fr * (txx1 * exx_el + tzz1 * ezz_el + (txx1 + tzz1) * (exx_el + ezz_el) + 2.0 * txz1 * exz_el);
            mesh->ddivpdexx_n[c0] += fr * ddivpdexx;
            mesh->ddivpdezz_n[c0] += fr * ddivpdezz;
            mesh->ddivpdgxz_n[c0] += fr * ddivpdexz / 2.0;
            mesh->ddivpdp_n[c0] += fr * ddivpdp;
            mesh->p_corr[c0] += fr * Pcorr;
            mesh->div_u_el[c0] += fr * div_el;
            mesh->div_u_pl[c0] += fr * div_pl;
            mesh->div_u_r[c0] += fr * div_r;
            if (UnsplitDiffReac == 0) mesh->X_n[c0] += fr * Xreac;
            mesh->OverS_n[c0] += fr * OverS;
            if (average == 1) {
                mesh->sxxd[c0] += fr * 1.0 / txx1;
                mesh->szzd[c0] += fr * 1.0 / tzz1;
                mesh->sxz_n[c0] += fr * 1.0 / txz1;
                mesh->eta_n[c0] += fr * 1.0 / etaVE;
                mesh->eta_phys_n[c0] += fr * 1.0 / eta;
                mesh->detadexx_n[c0] += fr * detadexx / pow(etaVE, 2.0);
                mesh->detadezz_n[c0] += fr * detadezz / pow(etaVE, 2.0);
                mesh->detadgxz_n[c0] += fr * detadexz / 2.0 / pow(etaVE, 2.0);
                mesh->detadp_n[c0] += fr * detadp / pow(etaVE, 2.0);
            }
            if (average == 2) {
                mesh->eta_n[c0] += fr * log(etaVE);
                mesh->eta_phys_n[c0] += fr * log(eta);
                mesh->detadexx_n[c0] += fr * detadexx / etaVE;
                mesh->detadezz_n[c0] += fr * detadezz / etaVE;
                mesh->detadgxz_n[c0] += fr * detadexz / 2.0 / etaVE;
                mesh->detadp_n[c0] += fr * detadp / etaVE;
            }
        }
        mesh->d_n[c0] = 1.0 / mesh->d_n[c0];
        if (average == 1) {
            mesh->sxxd[c0] = 1.0 / mesh->sxxd[c0];
            mesh->szzd[c0] = 1.0 / mesh->szzd[c0];
            mesh->sxz_n[c0] = 1.0 / mesh->sxz_n[c0];
            mesh->eta_n[c0] = 1.0 / mesh->eta_n[c0];
            mesh->eta_phys_n[c0] = 1.0 / mesh->eta_phys_n[c0];
            mesh->detadexx_n[c0] *= pow(mesh->eta_n[c0], 2.0);
            mesh->detadezz_n[c0] *= pow(mesh->eta_n[c0], 2.0);
            mesh->detadgxz_n[c0] *= pow(mesh->eta_n[c0], 2.0);
            mesh->detadp_n[c0] *= pow(mesh->eta_n[c0], 2.0);
            if (isinf(mesh->eta_phys_n[c0]) || isnan(mesh->eta_phys_n[c0])) {
                printf("%s: Problem on cell centers:\n", isinf(mesh->eta_phys_n[c0]) ? "Inf" : "NaN");
                exit(1);
            }
        }
        if (average == 2) {
            mesh->eta_n[c0] = exp(mesh->eta_n[c0]);
            mesh->eta_phys_n[c0] = exp(mesh->eta_phys_n[c0]);
            mesh->detadexx_n[c0] *= mesh->eta_n[c0];
            mesh->detadezz_n[c0] *= mesh->eta_n[c0];
            mesh->detadgxz_n[c0] *= mesh->eta_n[c0];
            mesh->detadp_n[c0] *= mesh->eta_n[c0];
        }
    }

#pragma omp parallel for schedule(dynamic, 256)
    for (int c1 = 0; c1 < Nx * Nz; c1++) {
        double eta = 0, txx1 = 0, tzz1 = 0, txz1 = 0, etaVE = 0, VEcoeff = 0, eII_el = 0, eII_pl = 0;
        double eII_pwl = 0, eII_exp = 0, eII_lin = 0, eII_gbs = 0, eII_cst = 0, dnew = 0;
        double exx_el = 0, ezz_el = 0, exz_el = 0, exx_diss = 0, ezz_diss = 0, exz_diss = 0;
        double detadexx = 0, detadezz = 0, detadexz = 0, detadp = 0, Xreac = 0, OverS = 0;
        double ddivpdexx = 0, ddivpdezz = 0, ddivpdexz = 0, ddivpdp = 0, Pcorr = 0, drhodp = 0, rho = 0;
        double div_el = 0, div_pl = 0, div_r = 0;
        mesh->VE_s[c1] = 0.0; mesh->sxz[c1] = 0.0; mesh->eta_phys_s[c1] = 0.0; mesh->eta_s[c1] = 0.0;
        mesh->exz_el[c1] = 0.0; mesh->exz_diss[c1] = 0.0;
        mesh->detadexx_s[c1] = 0.0; mesh->detadezz_s[c1] = 0.0; mesh->detadgxz_s[c1] = 0.0; mesh->detadp_s[c1] = 0.0;
        if (UnsplitDiffReac == 0) mesh->X_s[c1] = 0.0;
        mesh->OverS_s[c1] = 0.0;
        if (mesh->BCg.type[c1] == 30) continue;
        double etae = model->iselastic == 1 ? model->dt * mesh->mu_s[c1] : 1.0;
        double Exx, Ezz, Exz, Gxx, Gzz, Gxz;
        ve_strain_rates(model->aniso, model->aniso_fstrain, mesh->nx0_s[c1], mesh->nz0_s[c1],
                        mesh->aniso_factor_s[c1], mesh->FS_AR_s[c1], mesh->exxd_s[c1], mesh->ezzd_s[c1],
                        mesh->exz[c1], mesh->sxxd0_s[c1], mesh->szzd0_s[c1], mesh->sxz0[c1], etae,
                        &Exx, &Ezz, &Exz, &Gxx, &Gzz, &Gxz);
        for (int p = 0; p < model->Nb_phases; p++) {
            double fr = mesh->phase_perc_s[p][c1];
            int cond = fabs(fr) > 1.0e-13;
            if (!cond) continue;
            eta = orc_Viscosity(p, mesh->mu_s[c1], mesh->T_s[c1], mesh->P_s[c1], mesh->d0_s[c1], mesh->phi0_s[c1], mesh->X0_s[c1],
                                Exx, Ezz, Exz, Gxx, Gzz, Gxz, mesh->sxxd0_s[c1], mesh->szzd0_s[c1], mesh->sxz0[c1],
                                materials, model, scaling, &txx1, &tzz1, &txz1, &etaVE, &VEcoeff, &eII_el, &eII_pl,
                                &eII_pwl, &eII_exp, &eII_lin, &eII_gbs, &eII_cst, &exx_el, &ezz_el, &exz_el,
                                &exx_diss, &ezz_diss, &exz_diss, &dnew, mesh->strain_s[c1], mesh->dil_s[c1],
                                mesh->fric_s[c1], mesh->C_s[c1], &detadexx, &detadezz, &detadexz, &detadp,
                                mesh->p0_s[c1], &Xreac, &OverS, &ddivpdexx, &ddivpdezz, &ddivpdexz, &ddivpdp,
                                &Pcorr, &drhodp, &rho, mesh->bet_s[c1], mesh->div_u_s[c1], &div_el, &div_pl, &div_r);
            if (average == 0) {
                mesh->eta_s[c1] += fr * etaVE;
                mesh->eta_phys_s[c1] += fr * eta;
                mesh->detadexx_s[c1] += fr * detadexx;
                mesh->detadezz_s[c1] += fr * detadezz;
                mesh->detadgxz_s[c1] += fr * detadexz / 2.0;
                mesh->detadp_s[c1] += fr * detadp;
            }
            if (average == 0 || average == 2) mesh->sxz[c1] += fr * txz1;
            mesh->VE_s[c1] += fr * VEcoeff;
            mesh->exz_el[c1] += fr * exz_el;
            mesh->exz_diss[c1] += fr * exz_diss;
            mesh->OverS_s[c1] += fr * OverS;
            if (UnsplitDiffReac == 0) mesh->X_s[c1] += fr * Xreac;
            if (average == 1) {
                mesh->sxz[c1] += fr * 1.0 / txz1;
                mesh->eta_s[c1] += fr * 1.0 / etaVE;
                mesh->eta_phys_s[c1] += fr * 1.0 / eta;
                mesh->detadexx_s[c1] += fr * detadexx / pow(etaVE, 2.0);
                mesh->detadezz_s[c1] += fr * detadezz / pow(etaVE, 2.0);
                mesh->detadgxz_s[c1] += fr * detadexz / 2.0 / pow(etaVE, 2.0);
                mesh->detadp_s[c1] += fr * detadp / pow(etaVE, 2.0);
            }
            if (average == 2) {
                mesh->eta_s[c1] += fr * log(etaVE);
                mesh->eta_phys_s[c1] += fr * log(eta);
                mesh->detadexx_s[c1] += fr * detadexx / etaVE;
                mesh->detadezz_s[c1] += fr * detadezz / etaVE;
                mesh->detadgxz_s[c1] += fr * detadexz / 2.0 / etaVE;
                mesh->detadp_s[c1] += fr * detadp / etaVE;
            }
        }
        if (average == 1) {
            mesh->sxz[c1] = 1.0 / mesh->sxz[c1];
            mesh->eta_s[c1] = 1.0 / mesh->eta_s[c1];
            mesh->eta_phys_s[c1] = 1.0 / mesh->eta_phys_s[c1];
            mesh->detadexx_s[c1] *= pow(mesh->eta_s[c1], 2.0);
            mesh->detadezz_s[c1] *= pow(mesh->eta_s[c1], 2.0);
            mesh->detadgxz_s[c1] *= pow(mesh->eta_s[c1], 2.0);
            mesh->detadp_s[c1] *= pow(mesh->eta_s[c1], 2.0);
            if (isinf(mesh->eta_phys_s[c1]) || isnan(mesh->eta_phys_s[c1])) {
                printf("%s: Problem on cell vertices:\n", isinf(mesh->eta_phys_s[c1]) ? "Inf" : "Nan");
                exit(1);
            }
        }
        if (average == 2) {
            mesh->eta_s[c1] = exp(mesh->eta_s[c1]);
            mesh->eta_phys_s[c1] = exp(mesh->eta_phys_s[c1]);
            mesh->detadexx_s[c1] *= mesh->eta_s[c1];
            mesh->detadezz_s[c1] *= mesh->eta_s[c1];
            mesh->detadgxz_s[c1] *= mesh->eta_s[c1];
            mesh->detadp_s[c1] *= mesh->eta_s[c1];
        }
    }
}

/* Anisotropy factors of one node (RheologyDensity.c:737-748). */
static void aniso_d(const params *model, double nx, double nz, double af, double fs, double *ani,
                    double *d0, double *d1) {
    *ani = 0.0; *d0 = 0.0; *d1 = 0.0;
    if (model->aniso != 0) {
        if (model->aniso_fstrain == 0) *ani = 1.0 - 1.0 / af;
        if (model->aniso_fstrain == 1) *ani = 1.0 - 1.0 / fs;
        *d0 = 2.0 * pow(nx, 2.0) * pow(nz, 2.0);
        *d1 = nx * nz * (-pow(nx, 2.0) + pow(nz, 2.0));
    }
}

/* RheologyDensity.c:713 RheologicalOperators */
void orc_RheologicalOperators(grid *mesh, params *model, scale *scaling, int Jacobian) {
    (void)scaling;
    int Nx = mesh->Nx, Nz = mesh->Nz, Ncx = Nx - 1, Ncz = Nz - 1;
    double dt = model->dt;
    int comp = (int)model->compressible;
    int status = 0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Ncx * Ncz; k++) {
        if (mesh->BCp.type[k] == 30 || mesh->BCp.type[k] == 31) {
            mesh->D11_n[k] = 0.0; mesh->D12_n[k] = 0.0; mesh->D13_n[k] = 0.0; mesh->D14_n[k] = 0.0;
            mesh->D21_n[k] = 0.0; mesh->D22_n[k] = 0.0; mesh->D23_n[k] = 0.0; mesh->D24_n[k] = 0.0;
            continue;
        }
        double ani, d0, d1, eta = mesh->eta_n[k];
        aniso_d(model, mesh->nx0_n[k], mesh->nz0_n[k], mesh->aniso_factor_n[k], mesh->FS_AR_n[k], &ani, &d0, &d1);
        if (Jacobian == 0) {
            mesh->D11_n[k] = 2.0 * eta - 2.0 * ani * d0 * eta;
            mesh->D12_n[k] = 2.0 * ani * d0 * eta;
            mesh->D13_n[k] = 2.0 * ani * d1 * eta;
            mesh->D14_n[k] = 0.0;
            mesh->D21_n[k] = 2.0 * ani * d0 * eta;
            mesh->D22_n[k] = 2.0 * eta - 2.0 * ani * d0 * eta;
            mesh->D23_n[k] = -2.0 * ani * d1 * eta;
            mesh->D24_n[k] = 0.0;
        } else {
            double etae = model->iselastic == 1 ? model->dt * mesh->mu_n[k] : 1.0;
            double K = comp == 1 ? 1.0 / mesh->bet_n[k] : 0.0;
            double Exx, Ezz, Exz, Gxx, Gzz, Gxz;
            ve_strain_rates(model->aniso, model->aniso_fstrain, mesh->nx0_n[k], mesh->nz0_n[k],
                            mesh->aniso_factor_n[k], mesh->FS_AR_n[k], mesh->exxd[k], mesh->ezzd[k],
                            mesh->exz_n[k], mesh->sxxd0[k], mesh->szzd0[k], mesh->sxz0_n[k], etae,
                            &Exx, &Ezz, &Exz, &Gxx, &Gzz, &Gxz);
            mesh->D11_n[k] = 2.0 * eta - 2.0 * ani * d0 * eta + 2.0 * mesh->detadexx_n[k] * Gxx - K * dt * mesh->ddivpdexx_n[k];
            mesh->D12_n[k] = 2.0 * ani * d0 * eta + 2.0 * mesh->detadezz_n[k] * Gxx - K * dt * mesh->ddivpdezz_n[k];
            mesh->D13_n[k] = 2.0 * ani * d1 * eta + 2.0 * mesh->detadgxz_n[k] * Gxx - K * dt * mesh->ddivpdgxz_n[k];
            mesh->D14_n[k] = 2.0 * mesh->detadp_n[k] * Exx - K * dt * mesh->ddivpdp_n[k];
            mesh->D21_n[k] = 2.0 * ani * d0 * eta + 2.0 * mesh->detadexx_n[k] * Gzz - K * dt * mesh->ddivpdexx_n[k];
            mesh->D22_n[k] = 2.0 * eta - 2.0 * ani * d0 * eta + 2.0 * mesh->detadezz_n[k] * Gzz - K * dt * mesh->ddivpdezz_n[k];
            mesh->D23_n[k] = -2.0 * ani * d1 * eta + 2.0 * mesh->detadgxz_n[k] * Gzz - K * dt * mesh->ddivpdgxz_n[k];
            mesh->D24_n[k] = 2.0 * mesh->detadp_n[k] * Ezz - K * dt * mesh->ddivpdp_n[k];
        }
    }
#pragma omp parallel for schedule(static) reduction(| : status)
    for (int k = 0; k < Nx * Nz; k++) {
        if (mesh->BCg.type[k] == 30) {
            mesh->D31_s[k] = 0.0; mesh->D32_s[k] = 0.0; mesh->D33_s[k] = 0.0; mesh->D34_s[k] = 0.0;
        } else {
            double ani, d0, d1, eta = mesh->eta_s[k];
            aniso_d(model, mesh->nx0_s[k], mesh->nz0_s[k], mesh->aniso_factor_s[k], mesh->FS_AR_s[k], &ani, &d0, &d1);
            if (Jacobian == 0) {
                mesh->D31_s[k] = 2.0 * ani * d1 * eta;
                mesh->D32_s[k] = -2.0 * ani * d1 * eta;
                mesh->D33_s[k] = eta + 2.0 * ani * (d0 - 0.5) * eta;
                mesh->D34_s[k] = 0.0;
            } else {
                double etae = model->iselastic == 1 ? model->dt * mesh->mu_s[k] : 1.0;
                double Exx, Ezz, Exz, Gxx, Gzz, Gxz;
                ve_strain_rates(model->aniso, model->aniso_fstrain, mesh->nx0_s[k], mesh->nz0_s[k],
                                mesh->aniso_factor_s[k], mesh->FS_AR_s[k], mesh->exxd_s[k], mesh->ezzd_s[k],
                                mesh->exz[k], mesh->sxxd0_s[k], mesh->szzd0_s[k], mesh->sxz0[k], etae,
                                &Exx, &Ezz, &Exz, &Gxx, &Gzz, &Gxz);
                mesh->D31_s[k] = 2.0 * ani * d1 * eta + 2.0 * mesh->detadexx_s[k] * Gxz;
                mesh->D32_s[k] = -2.0 * ani * d1 * eta + 2.0 * mesh->detadezz_s[k] * Gxz;
                mesh->D33_s[k] = eta + 2.0 * ani * (d0 - 0.5) * eta + 2.0 * mesh->detadgxz_s[k] * Gxz;
                mesh->D34_s[k] = 2.0 * mesh->detadp_s[k] * Exz;
            }
        }
        if (Jacobian == 1) {
            if (isnan(mesh->D34_s[k])) status |= 1;
            if (isinf(mesh->D34_s[k])) status |= 2;
        }
    }
    if (status & 1) { printf("EXIT: D34 is NAN!\n"); exit(1); }
    if (status & 2) { printf("EXIT: D34 is INF!\n"); exit(1); }
}

/* StokesRoutines.c:922 EvaluateRHS */
void orc_EvaluateRHS(grid *mesh, params model, scale scaling, double RHO_REF) {
    (void)scaling; (void)RHO_REF;
    int Nx = mesh->Nx, NxVz = Nx + 1, Nz = mesh->Nz, NzVx = Nz + 1, Ncx = Nx - 1;
    double dx = mesh->dx, dz = mesh->dz;
    for (int l = 0; l < NzVx; l++)
        for (int k = 0; k < Nx; k++) {
            int c = k + l * Nx, c2 = k + (l - 1) * Ncx;
            double gx = model.gx;
            mesh->roger_x[c] = 0.0;
            if (model.polar == 1) {
                double x = mesh->xg_coord[k], z = mesh->zvx_coord[l], tet = atan(z / x);
                if (tet > 0.0) gx = model.gz * cos(tet);
                if (tet < 0.0) gx = -model.gz * cos(tet);
            }
            mesh->gx[c] = gx;
            if (l > 0 && l < NzVx - 1 && (mesh->BCu.type[c] == -1 || mesh->BCu.type[c] == 2 || mesh->BCu.type[c] == -2)) {
                int iPrW = c2 - 1, iPrE = c2, ixyN = c, ixyS = c - Nx, iVxS = c - Nx, iVxN = c + Nx;
                if (mesh->BCu.type[c] == -2) iPrW = c2 + Ncx - 1;
                double inW = 0, inE = 0, inS = 0, inN = 0, rhoVx = 0.0;
                if (mesh->BCp.type[iPrW] == -1) inW = 1.0;
                if (mesh->BCp.type[iPrE] == -1) inE = 1.0;
                if (mesh->BCg.type[ixyS] != 30 && mesh->BCu.type[iVxS] != 13) inS = 1.0;
                if (mesh->BCg.type[ixyN] != 30 && mesh->BCu.type[iVxN] != 13) inN = 1.0;
                if (model.polar == 0) rhoVx = 0.5 * (mesh->rho_s[ixyS] + mesh->rho_s[ixyN]);
                if (model.polar == 1) rhoVx = 0.5 * (mesh->rho_n[iPrW] + mesh->rho_n[iPrE]);
                mesh->roger_x[c] = -gx * rhoVx;
                if (model.iselastic == 1 && k > 0 && k < Nx - 1) {
                    if (inE) mesh->roger_x[c] -= 1.0 / dx * (mesh->eta_n[iPrE] / (mesh->mu_n[iPrE] * model.dt) * mesh->sxxd0[iPrE]);
                    if (inW) mesh->roger_x[c] -= 1.0 / dx * (-mesh->eta_n[iPrW] / (mesh->mu_n[iPrW] * model.dt) * mesh->sxxd0[iPrW]);
                    if (inN) mesh->roger_x[c] -= 1.0 / dz * (mesh->eta_s[ixyN] / (mesh->mu_s[ixyN] * model.dt) * mesh->sxz0[ixyN]);
                    if (inS) mesh->roger_x[c] -= 1.0 / dz * (-mesh->eta_s[ixyS] / (mesh->mu_s[ixyS] * model.dt) * mesh->sxz0[ixyS]);
                }
            }
            mesh->roger_x[c] = -mesh->roger_x[c];
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < NxVz; k++) {
            int c = k + l * NxVz, c1 = k + l * Nx - 1, c2 = k - 1 + (l - 1) * Ncx;
            double gz = model.gz;
            mesh->roger_z[c] = 0.0;
            if (model.polar == 1) {
                double x = mesh->xvz_coord[k], z = mesh->zg_coord[l], tet = atan(z / x);
                if (tet > 0.0) gz = model.gz * sin(tet);
                if (tet < 0.0) gz = -model.gz * sin(tet);
            }
            mesh->gz[c] = gz;
            if (k > 0 && k < NxVz - 1 && mesh->BCv.type[c] == -1) {
                int iPrS = c2, iPrN = c2 + Ncx, ixyE = c1 + 1, ixyW = c1, iVyW = c - 1, iVyE = c + 1;
                double inS = 0, inN = 0, inW = 0, inE = 0;
                if (mesh->BCp.type[iPrS] == -1) inS = 1.0;
                if (mesh->BCp.type[iPrN] == -1) inN = 1.0;
                if (mesh->BCg.type[ixyW] != 30 && mesh->BCv.type[iVyW] != 13) inW = 1.0;
                if (mesh->BCg.type[ixyE] != 30 && mesh->BCv.type[iVyE] != 13) inE = 1.0;
                double rhoVz = 0.5 * (mesh->rho_n[iPrS] + mesh->rho_n[iPrN]);
                mesh->roger_z[c] = -gz * rhoVz;
                if (model.iselastic == 1 && l > 0 && l < Nz - 1) {
                    if (inN) mesh->roger_z[c] -= 1.0 / dz * (mesh->eta_n[iPrN] / (mesh->mu_n[iPrN] * model.dt) * (mesh->szzd0[iPrN]));
                    if (inS) mesh->roger_z[c] -= 1.0 / dz * (-mesh->eta_n[iPrS] / (mesh->mu_n[iPrS] * model.dt) * (mesh->szzd0[iPrS]));
                    if (inE) mesh->roger_z[c] -= 1.0 / dx * (mesh->eta_s[ixyE] / (mesh->mu_s[ixyE] * model.dt) * mesh->sxz0[ixyE]);
                    if (inW) mesh->roger_z[c] -= 1.0 / dx * (-mesh->eta_s[ixyW] / (mesh->mu_s[ixyW] * model.dt) * mesh->sxz0[ixyW]);
                }
            }
            mesh->roger_z[c] = -mesh->roger_z[c];
        }
    for (int l = 0; l < Nz - 1; l++)
        for (int k = 0; k < Nx - 1; k++) {
            int c = k + l * (Nx - 1);
            if (mesh->BCp.type[c] != 30 && mesh->BCp.type[c] != 31) {
                mesh->rhs_t[c] = mesh->T[c];
                mesh->rhs_p[c] = 0.0;
                mesh->Qrho[c] = 0.0;
                if (model.compressible == 1 && mesh->comp_cells[c] == 1) {
                    if (model.VolChangeReac == 0) mesh->rhs_p[c] += mesh->p0_n[c] * mesh->bet_n[c] / model.dt;
                    if (model.VolChangeReac == 1) mesh->rhs_p[c] += log(mesh->rho0_n[c]) / model.dt;
                    if (model.adiab_heat > 0) mesh->rhs_p[c] += mesh->divth0_n[c];
                }
            }
        }
}
