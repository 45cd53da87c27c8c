"""MI355X-native lyotropic active-nematic step (D2Q9 LB + Q-tensor + phase
field) behind the reference's Model interface; see include/mass_lyo.h."""
