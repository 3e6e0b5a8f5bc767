"""MI355X-native metric-dependent CCH path (customize / negative-cycle scan /
partial update / elimination-tree query) behind the reference's Python entry
points."""
