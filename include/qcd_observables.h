/* qcd_observables.h -- gauge observables of the MI355X library's drop-in API.
 *
 * Counterpart of include/qcd_observables.h:12 (lib/src/qcd_observables.c:15-36).
 */
#ifndef H_QCD_OBSERVABLES
#define H_QCD_OBSERVABLES 1

#ifdef __cplusplus
extern "C" {
#endif

/* mean plaquette: sum over sites and the 6 planes mu < nu of
 * Re tr[U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+], divided by 18 V; on the GPU */
qcd_real_8 qcd_calculatePlaquette(qcd_gaugeField *u);

#ifdef __cplusplus
}
#endif

#endif
