"""MI355X-native backend for the specific_particle_tracer tracing path."""
