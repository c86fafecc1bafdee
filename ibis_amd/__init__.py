"""ibis_amd -- MI355X-native cell residual (dU/dt) for the ibis finite-volume solver."""
