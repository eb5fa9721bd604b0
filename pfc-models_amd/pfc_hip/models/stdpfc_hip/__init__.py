"""standard PFC on the hip backend (libpfc_hip.so on MI355X)."""
