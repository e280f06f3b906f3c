"""tsmi -- MI355X-native Trotter-Suzuki evolution step (host-side Python view)."""
