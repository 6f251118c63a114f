"""preclib_amd: MI355X-native large-integer product path for preclib's precn_t."""
