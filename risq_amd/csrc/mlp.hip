// Learned Davidson preconditioner (MLP).  Placeholder until the network lands.
#include "mlp.h"

bool mlp_ready(const hci_ctx* c) { return c && c->mlp; }
int mlp_correction(hci_ctx* c, const hci_csr*, double, const double*, const double*, const double*,
                   double*) {
    return hci_fail(c, HCI_ERR_STATE, "MLP preconditioner not loaded");
}
void mlp_free_state(hci_ctx*) {}
