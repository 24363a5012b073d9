"""Data objects on the SH hot path (mirrors specula/data_objects/)."""
