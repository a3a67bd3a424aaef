// Runs the reference program built by oracle/Makefile (test infrastructure):
// the CPU baseline executable is this file linked against a reference shim
// library of one configuration.
extern "C" int ref_main();
int main() { return ref_main(); }
