"""MI355X-native pyramid Lucas-Kanade registration (drop-in for getMotion / correctMotion)."""
