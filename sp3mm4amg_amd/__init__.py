"""MI355X-native Galerkin triple product R*A*P (Sp3MM) on sorted CSR."""
