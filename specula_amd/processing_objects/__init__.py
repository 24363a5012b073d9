"""Processing objects on the SH hot path (mirrors specula/processing_objects/)."""
