"""MI355X-native spin-projection grid kernel (PHFTools_F drop-in)."""
