"""helios_amd — MI355X-native path-tracing hot path of Helios (package root)."""
