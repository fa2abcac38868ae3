"""MI355X-native pressure solver for hmfem's cell-centred Darcy model."""
