"""laminarsmoke_amd: MI355X-native operator-splitting reaction step of laminarSMOKE."""
