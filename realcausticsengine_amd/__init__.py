"""MI355X-native vertex-merging caustics engine (drop-in for the Real Caustics C-ABI)."""
