// gle_newton.h -- Newton-Raphson branch-length optimisation as a resumable state
// machine, shared by the tree layer (one branch at a time) and the serving
// layer (one iteration per request per batch).
//
// Restates Tree::NewtonRaphsonOptimizeBranchLength, FP64 branch, without its
// debugging output (trunk/src/optimization.cpp:1359-1972): Newton step -d1/d2,
// known-min/known-max bracketing, fallbacks when d2 >= 0, exit when the
// estimated gain falls below the precision, at most 100 iterations, lengths in
// [1e-8, 100] (defs.h:106-107).
#pragma once
#include <algorithm>
#include <cmath>

namespace gle {

struct NewtonBranch {
  enum Action { kContinue, kFinish, kSetAndFinish };
  double prec = 0.0, min_b = 1e-8, max_b = 100.0;
  double total = 0.0, v = 0.0, v_prev = 0.0, knownMin = 1e-8, knownMax = 100.0, initialL = 0.0;
  int iter = 0, negProposalNum = 0;

  void start(double v0, double precision) {
    *this = NewtonBranch();
    prec = precision;
    v = v_prev = v0;
  }

  // Derivatives (lnL, d1, d2) at the branch's current length dlen.  On
  // kContinue/kSetAndFinish the caller sets the branch to *next.
  Action step(double lnL, double d1, double d2, double dlen, double *next) {
    if (iter == 0) initialL = lnL;
    const double estDeltaNR = -d1 / d2;
    const double estScoreDelta = d1 * estDeltaNR + (d2 * estDeltaNR * estDeltaNR * 0.5);
    if (d1 <= 0.0 && dlen < knownMax) knownMax = dlen;
    else if (d1 > 0.0 && dlen > knownMin) knownMin = dlen;
    if (d2 >= 0.0) {  // curvature wrong for a Newton step
      double proposed;
      if (d1 <= 0.0) {
        if (std::fabs(dlen - min_b) < 1.0e-8) return kFinish;
        if (knownMin == min_b) {
          if (dlen <= 1.0e-4) proposed = min_b;
          else if (dlen <= 0.005) proposed = 1.0e-4;
          else if (dlen <= 0.05) proposed = dlen * 0.1;
          else proposed = dlen * 0.25;
        } else
          proposed = (knownMin + dlen) * 0.5;
        if (iter > 0 || proposed == min_b) {
          const double estImp = d1 * (proposed - dlen) + (d2 * (proposed - dlen) * (proposed - dlen) * 0.5);
          if (estImp < prec && lnL >= initialL - 1.0e-4) return kFinish;
        }
      } else {
        if (knownMax == max_b) proposed = dlen < 0.1 ? dlen * 10.0 : std::min(dlen * 2.0, max_b);
        else proposed = (knownMax + dlen) * 0.5;
        if (iter > 0) {
          const double estImp = d1 * (proposed - dlen) + (d2 * (proposed - dlen) * (proposed - dlen) * 0.5);
          if (estImp < prec && lnL >= initialL - 1.0e-4) return kFinish;
        }
      }
      v = proposed;
      total += prec;
    } else {
      if (d1 < 0.0 && std::fabs(dlen - min_b) < 1.0e-8) return kFinish;
      if (d1 > 0.0 && std::fabs(dlen - max_b) < 1.0e-8) return kFinish;
      if (estScoreDelta < prec && (iter == 0 || lnL + ((iter < 10 ? 1 : iter) * 1.0e-7) >= initialL))
        return kFinish;
      bool noNR = false;
      if (iter > 10) {
        if (d1 > 0.0) {
          if (std::fabs(knownMax - max_b) < 1e-8) {
            if (iter > 20 && dlen > max_b / 2.0) v = max_b;
            else if (v < 0.2) v = std::min((v + max_b) * 0.5, v * 5.0);
            else v = std::min((v + max_b) * 0.5, v * 2.0);
            noNR = true;
          } else if ((iter - 20) > 0 && (iter - 20) % 10 == 0)
            v = std::min((v + knownMax) * 0.5, v * 100.0);
        } else if (std::fabs(knownMin - min_b) < 1e-8) {
          v = iter > 20 ? min_b : (v + min_b) * 0.5;
          noNR = true;
        }
      }
      if (!noNR) v += estDeltaNR;
      if (v <= knownMin) {
        negProposalNum++;
        if (knownMin == min_b) {
          const double dm = min_b - dlen, sdm = dm * d1 + (dm * dm * d2 * 0.5);
          if (sdm < prec) return kFinish;
          if (negProposalNum == 1 && dlen > 1e-4 && v_prev != 1e-4) {
            v = dlen < .005 ? 1e-4 : (dlen < 0.05 ? dlen * 0.1 : dlen * 0.25);
            const double delta = v - dlen;
            total += delta * d1 + (delta * delta * d2 * 0.5);
          } else {
            v = min_b;
            total += sdm;
          }
        } else {
          const double proposed = (knownMin + dlen) * 0.5, dm = proposed - dlen;
          const double sdm = dm * d1 + (dm * dm * d2 * 0.5);
          if (sdm < prec) return kFinish;
          v = proposed;
          total += sdm;
        }
      } else if (v >= knownMax) {
        if (knownMax == max_b) {
          const double dm = max_b - dlen, sdm = dm * d1 + (dm * dm * d2 * 0.5);
          if (sdm < prec) return kFinish;
          v = max_b;
          total += sdm;
        } else {
          const double proposed = (knownMax + dlen) * 0.5, dm = proposed - dlen;
          const double sdm = dm * d1 + (dm * dm * d2 * 0.5);
          if (sdm < prec) return kFinish;
          v = proposed;
          total += sdm;
        }
      } else
        total += estScoreDelta;
    }
    *next = v;
    v_prev = v;
    return ++iter > 100 ? kSetAndFinish : kContinue;
  }
};

}  // namespace gle
