"""MI355X-native force-and-integrate path of the TCP_allpair two-component plasma."""
