// Learned Davidson preconditioner (MLP) -- internal interface.
#pragma once
#include "hci_internal.h"

bool mlp_ready(const hci_ctx* c);
// t = correction for the current Davidson step.  lam: Ritz value; r: residual;
// x: Ritz vector; dpr: the diagonal correction r / (lam - H_ii); all length n.
int mlp_correction(hci_ctx* c, const hci_csr* h, double lam, const double* r, const double* x,
                   const double* dpr, double* t);
void mlp_free_state(hci_ctx* c);
