/* placeholder: Nedelec restatement added below */
