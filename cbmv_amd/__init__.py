"""MI355X-native CBMV cost-volume pipeline (matchers -> features -> RF/MLP -> post)."""
