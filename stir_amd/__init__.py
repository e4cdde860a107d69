"""MI355X-native ray-tracing-matrix projector pair and OSEM for STIR."""
