"""MI355X-native implicit backward-Euler cloth step (klon3d-sim's ImplizitCg path)."""
