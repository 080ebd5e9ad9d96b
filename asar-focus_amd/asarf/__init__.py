"""asarf: MI355X-native range-Doppler SAR focusser (Python binding of the C ABI in include/asarf.h)."""
