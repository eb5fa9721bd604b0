"""log PFC with an external potential and the M-shifted split (logpfc_extv) on the hip backend."""
