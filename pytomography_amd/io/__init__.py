"""Readers that bring simulated SPECT data to the accelerated path (io/)."""
