// Learned Davidson preconditioner (MLP) -- internal interface.
#pragma once
#include "hci_internal.h"

bool mlp_ready(const hci_ctx* c);
// t = g * dpr with g from the network.  lam: Ritz value; r: residual; x: Ritz
// vector; dpr: the diagonal correction r / (lam - H_pp); rnorm/tnorm their
// norms; all vectors in internal order, length n.
int mlp_correction(hci_ctx* c, const hci_csr* h, double lam, double rnorm, double tnorm, const double* r,
                   const double* x, const double* dpr, double* t);
// raw per-determinant features (n x HCI_MLP_NFEAT)
int mlp_features(hci_ctx* c, const hci_csr* h, double lam, double rnorm, double tnorm, const double* r,
                 const double* x, const double* dpr, float* X);
void mlp_free_state(hci_ctx* c);
