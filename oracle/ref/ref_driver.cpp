// ref_driver.cpp -- C-ABI wrapper that compiles the reference's own
// Metric.h, Camera.h and BlackHole.h (BlackTracer2/BlackTracer/ under the
// reference checkout) into oracle/_ref/libbtref.so.
//
// TEST INFRASTRUCTURE ONLY.  It exists so the oracle's restatement of
// Path A can be checked bit-for-bit against the reference's own code.
// Nothing here is restated: every call below lands in a reference
// function.  Grid.h, PSHOffsetTable.h, Viewer.h and StarProcessor.h are
// not compiled: they include cuda_runtime.h, which this image lacks, so
// integration_wrapper's initial conditions (Grid.h:524-547) are supplied
// by the caller.
#include "Metric.h"
#include "Camera.h"
#include "BlackHole.h"

extern "C" {

// BlackHole(a) sets the process-global spin (BlackHole.h:12-15);
// Camera(theta, phi, r, br, btheta, bphi) derives the pose (Camera.h:40-61).
// out = {speed, alpha, w, wbar, Delta, ro}
void btref_pose(double a, double theta, double phi, double r, double br,
                double btheta, double bphi, double* out) {
    BlackHole bh(a);
    Camera cam(theta, phi, r, br, btheta, bphi);
    out[0] = cam.speed;
    out[1] = cam.alpha;
    out[2] = cam.w;
    out[3] = cam.wbar;
    out[4] = cam.Delta;
    out[5] = cam.ro;
}

int btref_check_celest(double a, double pR, double r, double theta, double b, double q) {
    metric::setAngVel(a);
    return metric::checkCelest(pR, r, theta, b, q) ? 1 : 0;
}

int btref_check_bq(double a, double b, double q) {
    metric::setAngVel(a);
    return metric::checkB_Q(b, q) ? 1 : 0;
}

int btref_check_rup(double a, double r, double theta, double b, double q) {
    metric::setAngVel(a);
    return metric::checkRup(r, theta, b, q) ? 1 : 0;
}

double btref_b0(double a, double r0) {
    metric::setAngVel(a);
    return metric::_b0(r0);
}

double btref_q0(double a, double r0) {
    metric::setAngVel(a);
    return metric::_q0(r0);
}

double btref_calc_speed(double a, double r, double theta) {
    metric::setAngVel(a);
    return metric::calcSpeed(r, theta);
}

void btref_derivs(double a, const double* var, double* out, double b, double q) {
    metric::setAngVel(a);
    double v[5] = {var[0], var[1], var[2], var[3], var[4]};
    metric::derivs(v, out, b, q);
}

void btref_rkck(double a, const double* var, const double* dvdz, double h, double* out,
                double* err, double b, double q) {
    metric::setAngVel(a);
    double aks[25];
    double tmp[5];
    metric::rkck(var, dvdz, 5, h, out, err, b, q, aks, tmp);
}

// rkckIntegrate1 (Metric.h:307-324): integrate one escaping ray.
void btref_integrate(double a, double r, double theta, double phi, double pR, double b,
                     double q, double pTheta, double* thetaOut, double* phiOut, int* step) {
    metric::setAngVel(a);
    double th = 0, ph = 0;
    int st = 0;
    metric::rkckIntegrate1(r, theta, phi, pR, b, q, pTheta, th, ph, st);
    *thetaOut = th;
    *phiOut = ph;
    *step = st;
}

// The loop body of integration_wrapper (Grid.h:548-553) after the initial
// conditions, over a batch of rays: the a-priori capture test and
// rkckIntegrate1 per ray, spread over OpenMP threads.  The spin is the
// reference's process-global, set once before the threads start and only
// read by them.  Captured rays get (-1, -1) and step 0, as there.
void btref_trace_batch(double a, double r, double theta, double phi, int n, const double* pR,
                       const double* pTheta, const double* b, const double* q, double* thetaOut,
                       double* phiOut, int* step, int nthreads) {
    metric::setAngVel(a);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int i = 0; i < n; i++) {
        double th = -1, ph = -1;
        int st = 0;
        if (metric::checkCelest(pR[i], r, theta, b[i], q[i]))
            metric::rkckIntegrate1(r, theta, phi, pR[i], b[i], q[i], pTheta[i], th, ph, st);
        thetaOut[i] = th;
        phiOut[i] = ph;
        step[i] = st;
    }
}

void btref_wrap_to_pi(double* theta, double* phi) {
    double t = *theta, p = *phi;
    metric::wrapToPi(t, p);
    *theta = t;
    *phi = p;
}

void btref_wrap_to_pi_f(float* theta, float* phi) {
    float t = *theta, p = *phi;
    metric::wrapToPi(t, p);
    *theta = t;
    *phi = p;
}

}  // extern "C"
