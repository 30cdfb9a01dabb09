"""MI355X-native per-blastwave radiation hot path of the PyBlastAfterglow afterglow code."""
