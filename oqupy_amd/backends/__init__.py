"""MI355X engine backends of the TEMPO and PT-TEMPO computations."""
